// Device form of the parameter/coefficient bytecode.
//
// The plan's expressions (rust/src/lindblad/eval.rs:11-153 opcodes, compounds in
// evaluation order then the Hamiltonian coefficients) are rewritten for the GPU
// without changing any arithmetic:
//
//  * In a time-dependent expression, every time-independent subexpression of
//    two or more instructions is moved into a hidden compound slot that is
//    evaluated once per trajectory; the expression reads the slot instead.
//    The operations and their order are unchanged, so the values are too.
//  * `CONST c; POW` with a real constant c becomes POWC, which carries
//    cos(c*pi) and sin(c*pi): for a negative real base the polar form
//    x^c = |x|^c (cos(c*atan2(0, x)), sin(c*atan2(0, x))) then needs no
//    trigonometric call at run time.
//
// Time-independent expressions are left as they are (they run once).

#include "lb_program.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

namespace {

struct Node {
  DevInstr ins;
  std::vector<int> kids;
  bool is_static = true;
  int size = 1;
};

int arity(const DevInstr& q) {
  switch (q.op) {
    case 1: case 2: case 3: case 4: return 0;
    case 10: case 11: case 20: return 1;
    case 12: case 13: return q.argc;
    default: return 2;  // binary arithmetic and comparisons
  }
}

// Postfix -> tree; returns the root or -1.
int parse(const lb_plan_desc* d, int start, int end, std::vector<Node>& nodes) {
  std::vector<int> st;
  for (int p = start; p < end; ++p) {
    Node n;
    n.ins = DevInstr{d->instr_op[p], d->instr_index[p], d->instr_argc[p], d->instr_function[p],
                     d->instr_re[p], d->instr_im[p], 0.0};
    const int a = arity(n.ins);
    if ((int)st.size() < a) return -1;
    n.kids.assign(st.end() - a, st.end());
    st.resize(st.size() - a);
    nodes.push_back(n);
    st.push_back((int)nodes.size() - 1);
  }
  return st.size() == 1 ? st[0] : -1;
}

void emit(const std::vector<Node>& nodes, int id, std::vector<DevInstr>& out) {
  for (int k : nodes[id].kids) emit(nodes, k, out);
  out.push_back(nodes[id].ins);
}

int depth_of(const std::vector<DevInstr>& prog, size_t start) {
  int sp = 0, peak = 0;
  for (size_t p = start; p < prog.size(); ++p) {
    const DevInstr& q = prog[p];
    sp += 1 - arity(q);
    peak = std::max(peak, sp);
  }
  return peak;
}

}  // namespace

int lb_build_program(const lb_plan_desc* d, const std::vector<uint8_t>& dyn_slot,
                     const std::vector<uint8_t>& compound_dyn, const std::vector<uint8_t>& coef_dyn,
                     LbProgram* out, std::string& why) {
  const int n_expr = d->n_compounds + d->n_coefficients;
  int next_slot = d->n_slots;
  std::vector<DevInstr>& prog = out->instr;
  prog.clear();
  out->expr_start.clear();
  out->compound_slot.clear();
  out->compound_dyn.clear();
  out->coef_dyn.assign(coef_dyn.begin(), coef_dyn.end());
  out->max_depth = 0;

  std::vector<std::vector<DevInstr>> hidden_progs;  // pending hidden compounds
  std::vector<int> hidden_slots;
  auto flush_hidden = [&]() {
    for (size_t h = 0; h < hidden_progs.size(); ++h) {
      out->expr_start.push_back((int32_t)prog.size());
      const size_t s0 = prog.size();
      prog.insert(prog.end(), hidden_progs[h].begin(), hidden_progs[h].end());
      out->max_depth = std::max(out->max_depth, depth_of(prog, s0));
      out->compound_slot.push_back(hidden_slots[h]);
      out->compound_dyn.push_back(0);
    }
    hidden_progs.clear();
    hidden_slots.clear();
  };

  std::vector<std::vector<DevInstr>> coef_progs(d->n_coefficients);
  for (int e = 0; e < n_expr; ++e) {
    const bool is_coef = e >= d->n_compounds;
    const bool dynamic = is_coef ? coef_dyn[e - d->n_compounds] != 0 : compound_dyn[e] != 0;
    std::vector<Node> nodes;
    const int root = parse(d, d->expr_ptrs[e], d->expr_ptrs[e + 1], nodes);
    if (root < 0) {
      why = "malformed expression " + std::to_string(e);
      return -1;
    }
    std::vector<DevInstr> body;
    if (!dynamic) {
      emit(nodes, root, body);
    } else {
      // staticness and sizes (children precede parents in `nodes`)
      for (Node& n : nodes) {
        if (n.ins.op == 4) n.is_static = false;
        if (n.ins.op == 2 && dyn_slot[n.ins.index]) n.is_static = false;
        for (int k : n.kids) {
          n.is_static = n.is_static && nodes[k].is_static;
          n.size += nodes[k].size;
        }
      }
      std::function<void(int)> hoist = [&](int id) {
        Node& n = nodes[id];
        for (int& k : n.kids) {
          if (nodes[k].is_static && nodes[k].size >= 2) {
            std::vector<DevInstr> sub;
            emit(nodes, k, sub);
            hidden_progs.push_back(sub);
            hidden_slots.push_back(next_slot);
            Node leaf;
            leaf.ins = DevInstr{2, next_slot, 0, 0, 0.0, 0.0, 0.0};
            ++next_slot;
            nodes.push_back(leaf);
            k = (int)nodes.size() - 1;
          } else if (!nodes[k].is_static) {
            hoist(k);
          }
        }
        // CONST c ; POW  ->  POWC(c) with cos(c*pi), sin(c*pi)
        if (n.ins.op == 9 && n.kids.size() == 2) {
          const Node& ex = nodes[n.kids[1]];
          if (ex.ins.op == 1 && ex.ins.im == 0.0) {
            const double c = ex.ins.re;
            n.ins = DevInstr{20, 0, 0, 0, c, std::cos(c * M_PI), std::sin(c * M_PI)};
            n.kids.resize(1);
          }
        }
      };
      if (nodes[root].is_static) {
        emit(nodes, root, body);
      } else {
        hoist(root);
        emit(nodes, root, body);
      }
    }
    if (!is_coef) {
      flush_hidden();  // hidden parts of a compound go right before it
      out->expr_start.push_back((int32_t)prog.size());
      const size_t s0 = prog.size();
      prog.insert(prog.end(), body.begin(), body.end());
      out->max_depth = std::max(out->max_depth, depth_of(prog, s0));
      out->compound_slot.push_back(d->compound_slots[e]);
      out->compound_dyn.push_back(compound_dyn[e]);
    } else {
      coef_progs[e - d->n_compounds] = body;
    }
  }
  flush_hidden();  // hidden parts of the coefficients follow all compounds
  out->n_compounds = (int)out->compound_slot.size();
  for (int k = 0; k < d->n_coefficients; ++k) {
    out->expr_start.push_back((int32_t)prog.size());
    const size_t s0 = prog.size();
    prog.insert(prog.end(), coef_progs[k].begin(), coef_progs[k].end());
    out->max_depth = std::max(out->max_depth, depth_of(prog, s0));
  }
  out->expr_start.push_back((int32_t)prog.size());
  out->n_slots = next_slot;
  if (prog.empty()) prog.push_back(DevInstr{1, 0, 0, 0, 0.0, 0.0, 0.0});
  return 0;
}
