// Host side of the C ABI in include/centrex_lindblad.h: plan validation and
// layout, device upload, kernel selection and launches.
//
// Plan checks follow rust/src/lindblad/plan.rs:1069-1153; coefficient time
// dependence follows plan.rs:1235-1291.  The lane layout is described in
// lb_device.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "lb_device.h"

extern "C" int lb_dev_variant_count(void);
extern "C" int lb_dev_variant(int id, int* R, int* maxw, int* waves);
extern "C" void* lb_dev_solve_kernel(int id, int solver);
extern "C" void* lb_dev_rhs_kernel(int id);
extern "C" void* lb_dev_queue_reset_kernel(void);

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(LB_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

}  // namespace

struct lb_plan {
  DevPlan dev{};
  std::vector<void*> allocations;
  int device = 0;
  int32_t* queue = nullptr;
  int default_variant = -1, default_W = 0;
  int64_t info[16] = {0};
  int64_t lds_bytes = 0;
  std::vector<int32_t> slot_up_host;  // slot -> upper index
};

namespace {

template <typename T>
int upload(lb_plan* p, const std::vector<T>& host, const T** out) {
  size_t bytes = std::max<size_t>(host.size(), 1) * sizeof(T);
  void* d = nullptr;
  HIP_TRY(hipMalloc(&d, bytes));
  p->allocations.push_back(d);
  if (!host.empty()) HIP_TRY(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(d);
  return LB_OK;
}

// Simulated stack depth of one expression; rejects what the device machine
// does not evaluate (temps, tuple helpers, unknown ids).
int check_expression(const lb_plan_desc* d, int start, int end, int n_slots, std::string& why) {
  int sp = 0, peak = 0;
  for (int p = start; p < end; ++p) {
    int op = d->instr_op[p];
    switch (op) {
      case 1: case 2: case 4: sp += 1; break;
      case 5: case 6: case 7: case 8: case 9: sp -= 1; break;
      case 10: case 11: break;
      case 12:
        if (d->instr_argc[p] < 1 || d->instr_function[p] < 1 || d->instr_function[p] > 7) {
          why = "unknown builtin function";
          return -1;
        }
        sp -= d->instr_argc[p] - 1;
        break;
      case 13: {
        int f = d->instr_function[p];
        if (f == 8 || f == 10 || f == 14 || f == 16) {
          why = "tuple-valued helper " + std::to_string(f) + " is not evaluated on the GPU";
          return -2;
        }
        if (f < 1 || f > 15) { why = "unknown helper id"; return -1; }
        sp -= d->instr_argc[p] - 1;
        break;
      }
      case 14: case 15: case 16: case 17: case 18: case 19: sp -= 1; break;
      case 3: why = "TEMP instructions belong to entrywise plans"; return -2;
      default: why = "unknown instruction op " + std::to_string(op); return -1;
    }
    if (op == 2 && (d->instr_index[p] < 0 || d->instr_index[p] >= n_slots)) {
      why = "slot index out of range";
      return -1;
    }
    if (sp < 1) { why = "stack underflow"; return -1; }
    peak = std::max(peak, sp);
  }
  if (sp != 1) { why = "expression does not end with a single value"; return -1; }
  if (peak > LB_MAX_STACK) { why = "expression stack deeper than the device machine"; return -2; }
  return 0;
}

// Smallest workgroup (multiple of 64) with which variant `id` covers S slots, or 0.
int variant_threads(int id, int S) {
  int R, maxw, waves;
  if (!lb_dev_variant(id, &R, &maxw, &waves)) return 0;
  int W = ((S + R - 1) / R + 63) / 64 * 64;
  if (W < 64) W = 64;
  return W <= maxw ? W : 0;
}

// Default variant: the most registers per lane that still keeps two
// workgroups' worth of waves per SIMD pair, then the least padding.  The
// autotuner (centrex_tlf_amd.lindblad.autotune) replaces this by measurement.
void pick_config(int S, int* id_out, int* W_out) {
  int best = -1, bestW = 0;
  int64_t bestScore = INT64_MAX;
  for (int id = 0; id < lb_dev_variant_count(); ++id) {
    int R, maxw, waves;
    lb_dev_variant(id, &R, &maxw, &waves);
    int W = variant_threads(id, S);
    if (!W) continue;
    const int64_t waste = (int64_t)R * W - S;
    const int64_t score = waste + (waves == 2 ? 0 : 256) + 8 * R;
    if (score < bestScore) { bestScore = score; best = id; bestW = W; }
  }
  *id_out = best;
  *W_out = bestW;
}

// Must match carve() in lb_kernels.hip.
int64_t lds_bytes_for(const DevPlan& P) {
  return 8 * (int64_t)(4 * P.n_pub) + 8 * 2 * (int64_t)P.S + 8 * (int64_t)((P.n + 1) & ~1) +
         16 * 8 * (int64_t)(P.K + 1) + 16 * 8 * (int64_t)P.n_slots + 8 * 4 * (int64_t)P.nblk + 16;
}

int set_lds_limit(const void* fn, size_t lds) {
  if (lds > 65536)
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return LB_OK;
}

}  // namespace

extern "C" const char* lb_last_error(void) { return g_error.c_str(); }

extern "C" int lb_plan_create(const lb_plan_desc* d, lb_plan** out_plan) {
  if (!out_plan) return fail(LB_ERR_INVALID, "out_plan is null");
  *out_plan = nullptr;
  if (!d || d->n_states <= 0) return fail(LB_ERR_INVALID, "n_states must be positive");
  const int n = d->n_states, U = n * (n + 1) / 2;
  if (d->n_base < 0 || d->n_base > d->n_slots) return fail(LB_ERR_INVALID, "n_base exceeds n_slots");
  if (d->n_coefficients > 254) return fail(LB_ERR_UNSUPPORTED, "more than 254 Hamiltonian coefficients");
  const int n_expr = d->n_compounds + d->n_coefficients;
  for (int c = 0; c < d->n_compounds; ++c)
    if (d->compound_slots[c] < d->n_base || d->compound_slots[c] >= d->n_slots)
      return fail(LB_ERR_INVALID, "compound slot out of range");
  for (int e = 0; e < n_expr; ++e) {
    std::string why;
    int rc = check_expression(d, d->expr_ptrs[e], d->expr_ptrs[e + 1], d->n_slots, why);
    if (rc == -1) return fail(LB_ERR_INVALID, "expression " + std::to_string(e) + ": " + why);
    if (rc == -2) return fail(LB_ERR_UNSUPPORTED, "expression " + std::to_string(e) + ": " + why);
  }
  const int64_t T = d->n_terms;
  if (d->output_ptrs[0] != 0 || d->output_ptrs[U] != T)
    return fail(LB_ERR_INVALID, "expanded RHS output_ptrs must start at 0 and end at term count");
  for (int o = 0; o < U; ++o)
    if (d->output_ptrs[o] > d->output_ptrs[o + 1])
      return fail(LB_ERR_INVALID, "expanded RHS output_ptrs must be monotonic");
  for (int64_t k = 0; k < T; ++k) {
    if (d->input_indices[k] < 0 || d->input_indices[k] >= U)
      return fail(LB_ERR_INVALID, "expanded RHS input index out of bounds for upper length");
    if (d->coeff_indices[k] >= d->n_coefficients)
      return fail(LB_ERR_INVALID, "expanded RHS coefficient index out of bounds");
  }

  // ---- time dependence (plan.rs:1235-1291)
  std::vector<uint8_t> dyn_slot(d->n_slots, 0), compound_dyn(d->n_compounds, 0), coef_dyn(d->n_coefficients, 0);
  bool time_dependent = false;
  auto reads_dynamic = [&](int e, bool* uses_time) {
    bool dyn = false;
    *uses_time = false;
    for (int p = d->expr_ptrs[e]; p < d->expr_ptrs[e + 1]; ++p) {
      if (d->instr_op[p] == 4) { dyn = true; *uses_time = true; }
      if (d->instr_op[p] == 2 && dyn_slot[d->instr_index[p]]) dyn = true;
    }
    return dyn;
  };
  for (int c = 0; c < d->n_compounds; ++c) {
    bool ut;
    compound_dyn[c] = reads_dynamic(c, &ut);
    dyn_slot[d->compound_slots[c]] = compound_dyn[c];
    time_dependent |= ut;
  }
  int n_dyn_coef = 0;
  for (int k = 0; k < d->n_coefficients; ++k) {
    bool ut;
    coef_dyn[k] = reads_dynamic(d->n_compounds + k, &ut);
    time_dependent |= ut;
    n_dyn_coef += coef_dyn[k];
  }

  // ---- upper index -> packed indices (rhs.rs:151-176)
  std::vector<int32_t> up_re(U), up_im(U);
  {
    int cursor = n, u = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++u) {
        if (i == j) { up_re[u] = i; up_im[u] = -1; }
        else { up_re[u] = cursor; up_im[u] = cursor + 1; cursor += 2; }
      }
  }

  // ---- split terms into self (own element, time-independent coefficient) and gathers
  auto is_dyn = [&](int64_t k) { return d->coeff_indices[k] >= 0 && coef_dyn[d->coeff_indices[k]]; };
  std::vector<int> n_gather(U, 0), n_self(U, 0);
  std::vector<uint8_t> published(U, 0);
  int64_t flops = 0;
  for (int o = 0; o < U; ++o) {
    for (int64_t k = d->output_ptrs[o]; k < d->output_ptrs[o + 1]; ++k) {
      const int in = (int)d->input_indices[k];
      flops += (up_im[in] < 0) ? 4 : 8;
      if (is_dyn(k)) flops += 6;
      if (in == o && !is_dyn(k) && !d->input_conj[k]) {
        ++n_self[o];
      } else {
        ++n_gather[o];
        published[in] = 1;
      }
    }
  }
  // slots: outputs ordered by gather count (descending), then index
  std::vector<int> order(U);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n_gather[a] > n_gather[b]; });
  const int S = (U + 63) / 64 * 64, nblk = S / 64;
  std::vector<int32_t> slot_up(S, -1), slot_re(S, -1), slot_im(S, -1), slot_pub(S, -1);
  std::vector<int32_t> pub_of_upper(U, -1);
  int n_pub = 0;
  for (int s = 0; s < U; ++s) {
    const int o = order[s];
    slot_up[s] = o;
    slot_re[s] = up_re[o];
    slot_im[s] = up_im[o];
    if (published[o]) { pub_of_upper[o] = n_pub; slot_pub[s] = n_pub++; }
  }
  if (n_pub > (1 << 20)) return fail(LB_ERR_UNSUPPORTED, "too many published elements");
  int max_self = 1;
  for (int o = 0; o < U; ++o) max_self = std::max(max_self, n_self[o]);
  std::vector<int32_t> self_coef((size_t)S * max_self, -2);
  std::vector<uint8_t> self_conj((size_t)S * max_self, 0);
  std::vector<double> self_f((size_t)S * max_self * 2, 0.0);
  std::vector<int32_t> blk_row0(nblk), blk_rows(nblk);
  int rows_total = 0;
  for (int b = 0; b < nblk; ++b) {
    int rows = 0;
    for (int l = 0; l < 64; ++l) {
      int s = 64 * b + l;
      if (s < U) rows = std::max(rows, n_gather[order[s]]);
    }
    blk_row0[b] = rows_total;
    blk_rows[b] = rows;
    rows_total += rows;
  }
  std::vector<uint32_t> g_word((size_t)rows_total * 64, lb_gword(0, true, false, false, LB_STATIC_COEF));
  std::vector<double> g_f((size_t)rows_total * 64 * 2, 0.0);
  for (int s = 0; s < U; ++s) {
    const int o = order[s], b = s / 64, l = s % 64;
    int js = 0, jg = 0;
    for (int64_t k = d->output_ptrs[o]; k < d->output_ptrs[o + 1]; ++k) {
      const int in = (int)d->input_indices[k];
      const int64_t ci = d->coeff_indices[k];
      if (in == o && !is_dyn(k) && !d->input_conj[k]) {
        const size_t e = (size_t)s * max_self + js++;
        self_coef[e] = (int32_t)ci;
        self_conj[e] = d->coeff_conj[k];
        self_f[2 * e] = d->factors_re[k];
        self_f[2 * e + 1] = d->factors_im[k];
      } else {
        const size_t e = (size_t)(blk_row0[b] + jg++) * 64 + l;
        g_word[e] = lb_gword((uint32_t)pub_of_upper[in], up_im[in] < 0, d->input_conj[k] != 0,
                             d->coeff_conj[k] != 0, ci < 0 ? LB_STATIC_COEF : (uint32_t)ci);
        g_f[2 * e] = d->factors_re[k];
        g_f[2 * e + 1] = d->factors_im[k];
      }
    }
  }

  // ---- expression machine
  std::vector<DevInstr> instr(std::max(d->expr_ptrs[n_expr], 1));
  for (int p = 0; p < d->expr_ptrs[n_expr]; ++p)
    instr[p] = DevInstr{d->instr_op[p], d->instr_index[p], d->instr_argc[p], d->instr_function[p],
                        d->instr_re[p], d->instr_im[p]};
  std::vector<int32_t> expr_start(d->expr_ptrs, d->expr_ptrs + n_expr + 1);
  std::vector<int32_t> compound_slot(d->compound_slots, d->compound_slots + d->n_compounds);
  std::vector<double> base(2 * (size_t)d->n_base);
  for (int i = 0; i < d->n_base; ++i) { base[2 * i] = d->base_re[i]; base[2 * i + 1] = d->base_im[i]; }

  lb_plan* p = new lb_plan();
  (void)hipGetDevice(&p->device);
  DevPlan& P = p->dev;
  P.n = n;
  P.dim = n * n;
  P.upper_len = U;
  P.S = S;
  P.nblk = nblk;
  P.n_pub = n_pub;
  P.K = d->n_coefficients;
  P.n_slots = d->n_slots;
  P.n_base = d->n_base;
  P.n_compounds = d->n_compounds;
  P.max_self = max_self;
  P.time_dependent = time_dependent ? 1 : 0;
  int rc = LB_OK;
  auto up = [&](auto& v, auto** dst) {
    if (rc == LB_OK) rc = upload(p, v, dst);
  };
  up(slot_up, &P.slot_up);
  up(slot_re, &P.slot_re);
  up(slot_im, &P.slot_im);
  up(slot_pub, &P.slot_pub);
  up(self_coef, &P.self_coef);
  up(self_conj, &P.self_conj);
  up(self_f, &P.self_f);
  up(blk_row0, &P.blk_row0);
  up(blk_rows, &P.blk_rows);
  up(g_word, &P.g_word);
  up(g_f, &P.g_f);
  up(instr, &P.instr);
  up(expr_start, &P.expr_start);
  up(compound_slot, &P.compound_slot);
  up(compound_dyn, &P.compound_dyn);
  up(coef_dyn, &P.coef_dyn);
  up(base, &P.base);
  if (rc == LB_OK) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, 64);
    if (e != hipSuccess) rc = fail(LB_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    else { p->allocations.push_back(q); p->queue = static_cast<int32_t*>(q); }
  }
  if (rc != LB_OK) {
    lb_plan_destroy(p);
    return rc;
  }
  pick_config(S, &p->default_variant, &p->default_W);
  p->lds_bytes = lds_bytes_for(P);
  p->slot_up_host = slot_up;
  int64_t self_terms = 0, gather_terms = 0;
  for (int o = 0; o < U; ++o) { self_terms += n_self[o]; gather_terms += n_gather[o]; }
  int64_t* info = p->info;
  info[0] = n; info[1] = n * n; info[2] = U; info[3] = T; info[4] = d->n_coefficients;
  info[5] = time_dependent; info[6] = n_pub; info[7] = gather_terms; info[8] = self_terms;
  info[9] = S; info[10] = rows_total; info[11] = flops; info[12] = n_dyn_coef; info[13] = p->lds_bytes;
  info[14] = p->default_variant + 1; info[15] = p->default_W;
  *out_plan = p;
  return LB_OK;
}

extern "C" void lb_plan_destroy(lb_plan* p) {
  if (!p) return;
  for (void* a : p->allocations) (void)hipFree(a);
  delete p;
}

extern "C" int lb_plan_info(const lb_plan* p, int64_t* info) {
  if (!p || !info) return fail(LB_ERR_INVALID, "null argument");
  std::memcpy(info, p->info, sizeof(p->info));
  return LB_OK;
}

namespace {

// variant: 0 = plan default, k > 0 = kernel variant k - 1.  threads: 0 = the
// variant's smallest covering workgroup.
int resolve_config(const lb_plan* p, int variant, int threads, int* id, int* W) {
  if (threads < 0 || threads % 64 != 0)
    return fail(LB_ERR_INVALID, "threads_per_traj must be a multiple of 64");
  if (variant <= 0 && threads == 0) {
    *id = p->default_variant;
    *W = p->default_W;
    if (*id < 0) return fail(LB_ERR_UNSUPPORTED, "plan is too large for the available kernel variants");
    return LB_OK;
  }
  int R, maxw, waves;
  if (variant > 0) {
    *id = variant - 1;
    if (!lb_dev_variant(*id, &R, &maxw, &waves)) return fail(LB_ERR_INVALID, "unknown kernel variant");
    *W = threads ? threads : variant_threads(*id, p->dev.S);
    if (!*W || *W > maxw || (int64_t)R * *W < p->dev.S)
      return fail(LB_ERR_UNSUPPORTED, "kernel variant does not cover the plan with this workgroup size");
    return LB_OK;
  }
  // threads given: the variant with the fewest slots per thread that covers the plan
  for (int v = 0; v < lb_dev_variant_count(); ++v) {
    lb_dev_variant(v, &R, &maxw, &waves);
    if (threads <= maxw && (int64_t)R * threads >= p->dev.S) {
      *id = v;
      *W = threads;
      return LB_OK;
    }
  }
  return fail(LB_ERR_UNSUPPORTED, "no kernel variant covers the plan with this workgroup size");
}

int fill_overrides(const lb_plan* p, int32_t n, const int32_t* slots, int32_t* dst) {
  if (n < 0 || n > LB_MAX_OVERRIDES) return fail(LB_ERR_UNSUPPORTED, "at most 16 overridden parameters");
  for (int i = 0; i < n; ++i) {
    if (slots[i] < 0 || slots[i] >= p->dev.n_base)
      return fail(LB_ERR_INVALID, "parameter override slot " + std::to_string(slots[i]) +
                                      " is not a base parameter slot");
    dst[i] = slots[i];
  }
  return LB_OK;
}

}  // namespace

extern "C" int lb_rhs_eval(const lb_plan* p, int64_t batch, const double* y, const double* t,
                           int32_t n_override, const int32_t* override_slots,
                           const double* override_values, double* out, void* stream) {
  if (!p || !y || !t || !out) return fail(LB_ERR_INVALID, "null argument");
  if (batch <= 0) return LB_OK;
  DevRhsCall c{};
  c.batch = batch;
  c.y = y;
  c.t = t;
  int rc = fill_overrides(p, n_override, override_slots, c.override_slot);
  if (rc) return rc;
  c.n_override = n_override;
  c.override_values = override_values;
  c.out = out;
  int id, W;
  if ((rc = resolve_config(p, 0, 0, &id, &W))) return rc;
  void* fn = lb_dev_rhs_kernel(id);
  DevPlan P = p->dev;
  void* args[] = {&P, &c};
  if ((rc = set_lds_limit(fn, (size_t)p->lds_bytes))) return rc;
  int grid = (int)std::min<int64_t>(batch, 4096);
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, (size_t)p->lds_bytes, (hipStream_t)stream));
  HIP_TRY(hipGetLastError());
  return LB_OK;
}

extern "C" int lb_solve(const lb_plan* p, const lb_solve_args* a, void* stream) {
  if (!p || !a) return fail(LB_ERR_INVALID, "null argument");
  if (a->n_traj <= 0) return LB_OK;
  if (!a->y0 || !a->out_values || !a->out_stats || !a->out_fail_t)
    return fail(LB_ERR_INVALID, "y0 and output buffers are required");
  if (a->maxiters <= 0) return fail(LB_ERR_INVALID, "maxiters must be positive");
  if (a->solver != LB_SOLVER_DOPRI5 && a->solver != LB_SOLVER_TSIT5)
    return fail(LB_ERR_INVALID, "solver must be dopri5 or tsit5");
  if (!std::isfinite(a->t0) || (!a->t1_per_traj && !std::isfinite(a->t1)))
    return fail(LB_ERR_INVALID, "t_span must be finite");
  DevCall c{};
  c.n_traj = a->n_traj;
  c.t0 = a->t0;
  c.t1 = a->t1;
  c.t1_per_traj = a->t1_per_traj;
  c.abstol = a->abstol;
  c.reltol = a->reltol;
  c.dt = a->dt;
  c.maxiters = a->maxiters;
  c.save_start = a->save_start;
  c.solver = a->solver;
  c.y0 = a->y0;
  c.y0_shared = a->y0_shared;
  int rc = fill_overrides(p, a->n_override, a->override_slots, c.override_slot);
  if (rc) return rc;
  c.n_override = a->n_override;
  c.override_values = a->override_values;
  if (a->n_axes > 0) {
    if (a->n_axes != a->n_override || a->n_axes > LB_MAX_AXES)
      return fail(LB_ERR_INVALID, "grid axes must match the overridden slots (at most 8)");
    int64_t stride = 1, offset = 0, count = 1;
    for (int i = a->n_axes - 1; i >= 0; --i) {
      if (a->axis_lengths[i] <= 0) return fail(LB_ERR_INVALID, "parameter grid axes must be non-empty");
      c.axis_len[i] = a->axis_lengths[i];
      c.axis_stride[i] = stride;
      stride *= a->axis_lengths[i];
    }
    for (int i = 0; i < a->n_axes; ++i) { c.axis_offset[i] = offset; offset += a->axis_lengths[i]; count *= a->axis_lengths[i]; }
    if (count != a->n_traj) return fail(LB_ERR_INVALID, "n_traj must equal the grid size");
    c.n_axes = a->n_axes;
    c.axis_values = a->axis_values;
  } else if (a->n_override > 0 && !a->override_values) {
    return fail(LB_ERR_INVALID, "override_values are required");
  }
  c.output_kind = a->output_kind;
  if (a->output_kind < 0 || a->output_kind > 3) return fail(LB_ERR_INVALID, "unknown output kind");
  if (a->n_weights > LB_MAX_WEIGHTS) return fail(LB_ERR_UNSUPPORTED, "at most 64 integral weights");
  c.n_weights = a->n_weights;
  for (int i = 0; i < a->n_weights; ++i) {
    if (a->weight_indices[i] < 0 || a->weight_indices[i] >= p->dev.n)
      return fail(LB_ERR_INVALID, "integral weight index out of range");
    c.weight_index[i] = a->weight_indices[i];
    c.weight_value[i] = a->weight_values[i];
  }
  if (a->output_kind == LB_OUT_SELECTED) {
    if (a->n_selected > LB_MAX_SELECTED) return fail(LB_ERR_UNSUPPORTED, "at most 128 selected entries");
    const int n = p->dev.n;
    for (int q = 0; q < a->n_selected; ++q) {
      int i = a->selected_rows[q], j = a->selected_cols[q];
      if (i < 0 || j < 0 || i >= n || j >= n) return fail(LB_ERR_INVALID, "selected index out of range");
      int lo = std::min(i, j), hi = std::max(i, j);
      int u = lo * n - lo * (lo - 1) / 2 + (hi - lo);
      c.sel_upper[q] = u;
      c.sel_flags[q] = (i == j ? 1 : 0) | (i > j ? 2 : 0);
    }
    c.n_selected = a->n_selected;
  }
  c.out_values = a->out_values;
  c.out_stats = a->out_stats;
  c.out_fail_t = a->out_fail_t;
  c.out_integral = a->out_integral;
  c.want_integral = (a->output_kind == LB_OUT_INTEGRAL || a->out_integral != nullptr) ? 1 : 0;
  if (c.want_integral && a->n_weights <= 0)
    return fail(LB_ERR_INVALID, "integral outputs require integral weights");
  c.queue = p->queue;
  int id, W;
  if ((rc = resolve_config(p, a->variant, a->threads_per_traj, &id, &W))) return rc;
  void* fn = lb_dev_solve_kernel(id, a->solver);
  void* reset = lb_dev_queue_reset_kernel();
  hipStream_t s = (hipStream_t)stream;
  DevPlan P = p->dev;
  const size_t lds = (size_t)p->lds_bytes;
  if ((rc = set_lds_limit(fn, lds))) return rc;
  if (a->mode == LB_MODE_PLAIN) {
    // one launch per trajectory: the per-trajectory GPU path
    const int64_t total = a->n_traj;
    for (int64_t i = 0; i < total; ++i) {
      c.n_traj = 1;
      c.traj_offset = i;
      void* rargs[] = {&c.queue};
      HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
      void* args[] = {&P, &c};
      HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(W), args, lds, s));
    }
    HIP_TRY(hipGetLastError());
    return LB_OK;
  }
  int grid = a->workgroups;
  if (grid <= 0) {
    int per_cu = 0, cus = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, p->device));
    cus = prop.multiProcessorCount;
    grid = std::max(1, per_cu) * std::max(1, cus);
  }
  grid = (int)std::min<int64_t>(grid, a->n_traj);
  c.traj_offset = 0;
  void* rargs[] = {&c.queue};
  HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
  void* args[] = {&P, &c};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
  HIP_TRY(hipGetLastError());
  return LB_OK;
}

extern "C" int lb_variant_info(int32_t id, int32_t* slots_per_thread, int32_t* max_threads,
                               int32_t* waves_per_simd) {
  int R, maxw, waves;
  if (!lb_dev_variant(id, &R, &maxw, &waves)) return fail(LB_ERR_INVALID, "no such kernel variant");
  if (slots_per_thread) *slots_per_thread = R;
  if (max_threads) *max_threads = maxw;
  if (waves_per_simd) *waves_per_simd = waves;
  return LB_OK;
}
