// Device form of the parameter/coefficient bytecode.
//
// The plan's expressions (rust/src/lindblad/eval.rs:11-153 opcodes, compounds in
// evaluation order then the Hamiltonian coefficients) are rewritten for the GPU
// without changing any arithmetic:
//
//  * In a time-dependent expression, every time-independent subexpression of
//    two or more instructions is moved into a hidden compound slot that is
//    evaluated once per trajectory; the expression reads the slot instead.
//    Identical subexpressions share one slot.
//  * A time-dependent subexpression of two or more instructions that occurs
//    more than once across the plan (the same envelope for two lasers, a
//    polarisation switch and its complement, a phase factor used by two
//    couplings) is evaluated once per stage time: the first compound whose
//    whole expression it is keeps it, otherwise it moves into a hidden
//    time-dependent compound evaluated right before its first user, and every
//    occurrence reads the slot.
//    Each value is still produced by the same operations on the same operands,
//    so the results are bitwise those of the unshared program.
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
#include <map>
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

// Exact identity of a subtree: every field of every instruction, postfix order.
std::string key_of(const std::vector<Node>& nodes, int id) {
  std::string k;
  std::function<void(int)> walk = [&](int i) {
    for (int c : nodes[i].kids) walk(c);
    const DevInstr& q = nodes[i].ins;
    k.append(reinterpret_cast<const char*>(&q.op), sizeof q.op);
    k.append(reinterpret_cast<const char*>(&q.index), sizeof q.index);
    k.append(reinterpret_cast<const char*>(&q.argc), sizeof q.argc);
    k.append(reinterpret_cast<const char*>(&q.function), sizeof q.function);
    k.append(reinterpret_cast<const char*>(&q.re), sizeof q.re);
    k.append(reinterpret_cast<const char*>(&q.im), sizeof q.im);
    k.append(reinterpret_cast<const char*>(&q.aux), sizeof q.aux);
  };
  walk(id);
  return k;
}

Node leaf(int slot) {
  Node n;
  n.ins = DevInstr{2, slot, 0, 0, 0.0, 0.0, 0.0};
  return n;
}

// Subtree sizes (children precede parents in `nodes`).
void resize_tree(std::vector<Node>& nodes) {
  std::function<int(int)> size = [&](int i) {
    int s = 1;
    for (int c : nodes[i].kids) s += size(c);
    return nodes[i].size = s;
  };
  for (size_t i = 0; i < nodes.size(); ++i) size((int)i);
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

  // A hidden compound: program, slot, time-dependent.
  struct Hidden {
    std::vector<DevInstr> prog;
    int slot;
    bool dyn;
  };
  struct Expr {
    std::vector<Node> nodes;
    int root = -1;
    bool dynamic = false;
    std::vector<Hidden> before;  // hidden compounds evaluated before this expression
  };
  std::vector<Expr> ex(n_expr);

  // ---- phase A: parse; hoist time-independent subtrees (shared by key); POWC
  std::map<std::string, int> static_slot;
  for (int e = 0; e < n_expr; ++e) {
    Expr& X = ex[e];
    const bool is_coef = e >= d->n_compounds;
    X.dynamic = is_coef ? coef_dyn[e - d->n_compounds] != 0 : compound_dyn[e] != 0;
    X.root = parse(d, d->expr_ptrs[e], d->expr_ptrs[e + 1], X.nodes);
    if (X.root < 0) {
      why = "malformed expression " + std::to_string(e);
      return -1;
    }
    if (!X.dynamic) continue;
    std::vector<Node>& nodes = X.nodes;
    for (Node& n : nodes) {  // children precede parents
      n.is_static = !(n.ins.op == 4 || (n.ins.op == 2 && dyn_slot[n.ins.index]));
      for (int k : n.kids) n.is_static = n.is_static && nodes[k].is_static;
    }
    resize_tree(nodes);
    std::function<void(int)> hoist = [&](int id) {
      for (size_t c = 0; c < nodes[id].kids.size(); ++c) {
        const int k = nodes[id].kids[c];
        if (nodes[k].is_static && nodes[k].size >= 2) {
          const std::string key = key_of(nodes, k);
          auto it = static_slot.find(key);
          int slot;
          if (it != static_slot.end()) {
            slot = it->second;
          } else {
            std::vector<DevInstr> sub;
            emit(nodes, k, sub);
            slot = next_slot++;
            static_slot[key] = slot;
            X.before.push_back(Hidden{sub, slot, false});
          }
          nodes.push_back(leaf(slot));
          nodes[id].kids[c] = (int)nodes.size() - 1;
        } else if (!nodes[k].is_static) {
          hoist(k);
        }
      }
      // CONST c ; POW  ->  POWC(c) with cos(c*pi), sin(c*pi)
      Node& n = nodes[id];
      if (n.ins.op == 9 && n.kids.size() == 2) {
        const Node& e2 = nodes[n.kids[1]];
        if (e2.ins.op == 1 && e2.ins.im == 0.0) {
          const double c = e2.ins.re;
          n.ins = DevInstr{20, 0, 0, 0, c, std::cos(c * M_PI), std::sin(c * M_PI)};
          n.kids.resize(1);
        }
      }
    };
    if (!nodes[X.root].is_static) hoist(X.root);
    resize_tree(nodes);
  }

  // ---- phase B: occurrences of every time-dependent subtree of size >= 2
  std::map<std::string, int> count;
  std::function<void(const std::vector<Node>&, int, std::map<std::string, int>&, int)> tally =
      [&](const std::vector<Node>& nodes, int id, std::map<std::string, int>& into, int w) {
        const Node& n = nodes[id];
        if (n.is_static || n.size < 2) return;
        into[key_of(nodes, id)] += w;
        for (int k : n.kids) tally(nodes, k, into, w);
      };
  for (Expr& X : ex)
    if (X.dynamic && !X.nodes[X.root].is_static) tally(X.nodes, X.root, count, 1);

  // ---- phase C: share repeated time-dependent subtrees, emit
  std::map<std::string, int> dyn_slot_of;
  auto emit_hidden = [&](const Hidden& h) {
    out->expr_start.push_back((int32_t)prog.size());
    const size_t s0 = prog.size();
    prog.insert(prog.end(), h.prog.begin(), h.prog.end());
    out->max_depth = std::max(out->max_depth, depth_of(prog, s0));
    out->compound_slot.push_back(h.slot);
    out->compound_dyn.push_back(h.dyn ? 1 : 0);
  };
  std::vector<Hidden> coef_hidden;
  std::vector<std::vector<DevInstr>> coef_progs(d->n_coefficients);
  for (int e = 0; e < n_expr; ++e) {
    Expr& X = ex[e];
    const bool is_coef = e >= d->n_compounds;
    std::vector<Node>& nodes = X.nodes;
    std::vector<Hidden> shared;
    std::function<void(int, bool)> share = [&](int id, bool whole) {
      Node& n = nodes[id];
      if (n.is_static || n.size < 2) return;
      const std::string key = key_of(nodes, id);
      if (count[key] >= 2) {
        auto it = dyn_slot_of.find(key);
        if (it != dyn_slot_of.end()) {  // computed before: read its slot
          nodes[id] = leaf(it->second);
          nodes[id].is_static = false;
          return;
        }
        // the other occurrences will read a slot: their subtrees no longer count
        std::map<std::string, int> inner;
        for (int k : n.kids) tally(nodes, k, inner, count[key] - 1);
        for (auto& kv : inner) count[kv.first] -= kv.second;
        if (whole) {  // this compound's own slot holds the value
          dyn_slot_of[key] = d->compound_slots[e];
          for (int k : nodes[id].kids) share(k, false);
          return;
        }
        for (size_t c = 0; c < nodes[id].kids.size(); ++c) share(nodes[id].kids[c], false);
        Hidden h{{}, next_slot++, true};
        emit(nodes, id, h.prog);
        dyn_slot_of[key] = h.slot;
        shared.push_back(h);
        nodes[id] = leaf(h.slot);
        nodes[id].is_static = false;
        return;
      }
      for (size_t c = 0; c < nodes[id].kids.size(); ++c) share(nodes[id].kids[c], false);
    };
    if (X.dynamic && !nodes[X.root].is_static) share(X.root, !is_coef);
    std::vector<DevInstr> body;
    emit(nodes, X.root, body);
    std::vector<Hidden> pre = X.before;
    pre.insert(pre.end(), shared.begin(), shared.end());
    if (!is_coef) {
      for (const Hidden& h : pre) emit_hidden(h);  // right before the compound
      out->expr_start.push_back((int32_t)prog.size());
      const size_t s0 = prog.size();
      prog.insert(prog.end(), body.begin(), body.end());
      out->max_depth = std::max(out->max_depth, depth_of(prog, s0));
      out->compound_slot.push_back(d->compound_slots[e]);
      out->compound_dyn.push_back(compound_dyn[e]);
    } else {
      coef_hidden.insert(coef_hidden.end(), pre.begin(), pre.end());
      coef_progs[e - d->n_compounds] = body;
    }
  }
  for (const Hidden& h : coef_hidden) emit_hidden(h);  // after all compounds
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
