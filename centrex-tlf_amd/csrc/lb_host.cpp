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
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "lb_device.h"
#include "lb_internal.h"
#include "lb_program.h"
#include "lb_rtc.h"

extern "C" int lb_dev_variant_count(void);
extern "C" int lb_dev_variant(int id, int* R, int* maxw, int* waves);
extern "C" int lb_dev_variant_rows(int id);
extern "C" int lb_dev_variant_pair(int id);
extern "C" int lb_dev_adjoint_variant(int id, int* R, int* maxw, int* rows);
extern "C" int lb_dev_fixed_variant(int id, int* R, int* maxw);
extern "C" void* lb_dev_fixed_kernel(int id, int solver);
extern "C" int lb_dev_event_variant(int id, int* R, int* maxw);
extern "C" void* lb_dev_event_kernel(int id, int solver);
extern "C" void* lb_dev_adjoint_kernel(int id, int solver);
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

int is_capturing(hipStream_t s, bool* capturing) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(s, &cs));
  *capturing = cs != hipStreamCaptureStatusNone;
  return LB_OK;
}

}  // namespace

// Work-queue counters of a plan: every call takes its own counter, round
// robin from a ring, so solves of one plan on different streams or threads
// never share one.  Before a call resets its counter, its stream waits (on
// the device, by stream order) for the earlier call that used the same
// counter.  Inside a stream capture nothing is waited for or recorded: a
// captured call keeps its counter, and a replay of that graph must not
// overlap the call that reuses the counter kSlots calls later.
class QueueRing {
 public:
  static constexpr int kSlots = 64;
  static constexpr int kStride = 64;  // counters 256 bytes apart
  ~QueueRing() {
    for (hipEvent_t e : done_)
      if (e) (void)hipEventDestroy(e);
    if (base_) (void)hipFree(base_);
  }
  int init() {
    HIP_TRY(hipMalloc(&base_, sizeof(int32_t) * kSlots * kStride));
    HIP_TRY(hipMemset(base_, 0, sizeof(int32_t) * kSlots * kStride));
    return LB_OK;
  }
  // The counter of a new call on `s` (and its slot, for retire()).
  int acquire(hipStream_t s, int32_t** counter, int* slot) {
    bool cap = false;
    int rc = is_capturing(s, &cap);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(mu_);
    const int k = (int)(next_++ % kSlots);
    if (!cap && done_[k]) HIP_TRY(hipStreamWaitEvent(s, done_[k], 0));
    *counter = base_ + (size_t)k * kStride;
    *slot = k;
    return LB_OK;
  }
  // Mark the end of the call's launches on `s`.
  int retire(int slot, hipStream_t s) {
    bool cap = false;
    int rc = is_capturing(s, &cap);
    if (rc || cap) return rc;
    std::lock_guard<std::mutex> lock(mu_);
    if (!done_[slot]) HIP_TRY(hipEventCreateWithFlags(&done_[slot], hipEventDisableTiming));
    HIP_TRY(hipEventRecord(done_[slot], s));
    return LB_OK;
  }

 private:
  int32_t* base_ = nullptr;
  hipEvent_t done_[kSlots] = {};
  uint64_t next_ = 0;
  std::mutex mu_;
};

// Validated copy of a plan description plus the lowered expression program.
struct PlanSource {
  int n = 0, U = 0, K = 0, n_base = 0;
  int n_slots = 0;  // slots of the description (the program adds hidden ones after them)
  int n_diff = 0;  // derivative columns (parameters, then time when diff_time)
  int diff_time = 0;  // 1: the last derivative column is d/dt
  int64_t T = 0;
  std::vector<int64_t> output_ptrs, input_indices, coeff_indices;
  std::vector<uint8_t> coeff_conj, input_conj, coef_dyn;
  std::vector<double> factors_re, factors_im, base;
  std::vector<int32_t> up_re, up_im;  // upper index -> packed indices (-1 imag on the diagonal)
  LbProgram prog;
  std::vector<int32_t> expr_dyn;
  std::vector<uint8_t> fwd_compound;  // 0: a compound only the derivative programs read
  std::vector<int32_t> compound_slot;
  bool time_dependent = false;
  int n_dyn_coef = 0;
};

// Device tables for one set of active upper elements (all of them, or the
// closure of an initial-state support, see active_closure()).
struct DevImage {
  DevPlan dev{};
  std::vector<void*> allocations;
  int default_variant = -1, default_W = 0;
  int pair_variant = -1;  // paired variant for eligible calls (plans of at most 32 elements), or -1
  int64_t lds_bytes = 0;
  int64_t n_pub = 0, gather_terms = 0, self_terms = 0, rows_total = 0, flops = 0, active = 0;
  std::vector<uint8_t> active_mask;  // upper_len
  std::map<std::pair<int, int>, int> per_cu;  // (variant, W) -> resident workgroups per CU
  void* ws_y = nullptr;  // adjoint workspace (grow-only)
  void* ws_c = nullptr;
  size_t ws_y_len = 0, ws_c_len = 0;
  hipEvent_t ws_done = nullptr;  // end of the last adjoint that used the workspace
  std::vector<int32_t> g_rows_host;  // gather rows per block
  int t_rows_host = 0;               // transposed rows of block 0 (one-block plans)
};

struct lb_plan {
  PlanSource src;
  int device = 0, cus = 0;
  QueueRing queues;
  std::unique_ptr<DevImage> full;
  std::map<std::vector<uint8_t>, std::unique_ptr<DevImage>> restricted;
  std::mutex mu;  // guards `restricted`, the occupancy caches and the event buffers
  int64_t info[16] = {0};
  // stop-event program of the latest call: device copy, pinned staging, and
  // an event marking the end of the staging copy
  void* ev_dev = nullptr;
  void* ev_pinned = nullptr;
  size_t ev_cap = 0;
  hipEvent_t ev_copied = nullptr;
  // plan-specialised paired kernels (lb_plan_specialize): the generated
  // coefficient function, and (variant | gather rows << 8, solver) -> compiled module
  // A failed compilation leaves a null entry (the precompiled kernel runs)
  // and its reason in spec_failure.
  bool spec = false;
  std::string spec_source, spec_failure;
  std::map<std::pair<int, int>, std::pair<hipModule_t, hipFunction_t>> spec_kernels;
  ~lb_plan() {
    for (auto& kv : spec_kernels)
      if (kv.second.first) (void)hipModuleUnload(kv.second.first);
    auto release = [](DevImage* im) {
      if (!im) return;
      for (void* a : im->allocations) (void)hipFree(a);
      if (im->ws_y) (void)hipFree(im->ws_y);
      if (im->ws_c) (void)hipFree(im->ws_c);
      if (im->ws_done) (void)hipEventDestroy(im->ws_done);
    };
    release(full.get());
    for (auto& kv : restricted) release(kv.second.get());
    if (ev_copied) (void)hipEventSynchronize(ev_copied);
    if (ev_dev) (void)hipFree(ev_dev);
    if (ev_pinned) (void)hipHostFree(ev_pinned);
    if (ev_copied) (void)hipEventDestroy(ev_copied);
  }
};

namespace {

template <typename T>
int upload(DevImage* im, const std::vector<T>& host, const T** out) {
  size_t bytes = std::max<size_t>(host.size(), 1) * sizeof(T);
  void* d = nullptr;
  HIP_TRY(hipMalloc(&d, bytes));
  im->allocations.push_back(d);
  if (!host.empty()) HIP_TRY(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(d);
  return LB_OK;
}

// Simulated stack depth of one expression; rejects what the device machine
// does not evaluate (temps, tuple helpers, unknown ids).
int check_expression(const lb_plan_desc* d, int start, int end, int n_slots, std::string& why,
                     int* depth = nullptr) {
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
        const int f = d->instr_function[p], n = d->instr_argc[p];
        if (f < 1 || f > 17) { why = "unknown helper id"; return -1; }
        if (n < 1) { why = "helper without arguments"; return -1; }
        // tuple helpers take their tuples' elements as consecutive arguments
        bool shape = true;
        if (f == 8) shape = n >= 4 && (n - 4) % 3 == 0;
        if (f == 10) shape = n >= 5 && ((n - 5) % 3 == 0 || (n - 6) % 3 == 0);
        if (f == 14 || f == 16) shape = n >= 3 && n % 2 == 1;
        if (f == 17) shape = n >= 5 && n % 2 == 1;
        if (!shape) {
          why = "helper " + std::to_string(f) + " with " + std::to_string(n) + " arguments";
          return -1;
        }
        sp -= n - 1;
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
  if (depth) *depth = peak;
  return 0;
}

// Smallest workgroup (multiple of 64) with which variant `id` covers S slots
// and `rows` gather rows, or 0.
int variant_threads(int id, int S, int rows) {
  int R, maxw, waves;
  if (!lb_dev_variant(id, &R, &maxw, &waves)) return 0;
  if (const int g = lb_dev_variant_rows(id)) return (S <= 64 && rows <= g) ? 64 : 0;
  int W = ((S + R - 1) / R + 63) / 64 * 64;
  if (W < 64) W = 64;
  return W <= maxw ? W : 0;
}

// Default variant: the most registers per lane that still keeps two
// workgroups' worth of waves per SIMD pair, then the least padding.  The
// autotuner (centrex_tlf_amd.lindblad.autotune) replaces this by measurement.
void pick_config(int S, int rows, int* id_out, int* W_out) {
  int best = -1, bestW = 0;
  int64_t bestScore = INT64_MAX;
  for (int id = 0; id < lb_dev_variant_count(); ++id) {
    int R, maxw, waves;
    lb_dev_variant(id, &R, &maxw, &waves);
    if (lb_dev_variant_pair(id)) continue;  // chosen per call (lb_solve)
    int W = variant_threads(id, S, rows);
    if (!W) continue;
    const int64_t waste = (int64_t)R * W - S;
    const int g = lb_dev_variant_rows(id);
    // single-wave variants first: the fewest register rows that hold the
    // plan, at two waves per SIMD (256 registers per lane, no spills)
    const int64_t score = g ? -1000 + g + (waves == 2 ? 0 : 1)
                            : waste + (waves == 2 ? 0 : 256) + 8 * R;
    if (score < bestScore) { bestScore = score; best = id; bestW = W; }
  }
  *id_out = best;
  *W_out = bestW;
}

// Gather tables up to this size are copied into LDS by every workgroup.
constexpr int64_t kLdsTableBytes = 32 * 1024;
// LDS of one workgroup on gfx950.
constexpr int64_t kMaxLdsBytes = 160 * 1024;

// Must match carve() in lb_kernels.hip.
int64_t lds_bytes_for(const DevPlan& P) {
  const int64_t cells = P.lds_tables ? (int64_t)P.g_rows * 64 : 0;
  return 8 * (int64_t)(4 * P.n_pub) + 8 * 2 * (int64_t)P.S + 8 * (int64_t)((P.n + 1) & ~1) +
         16 * 8 * (int64_t)(P.K + 1) + 16 * 8 * (int64_t)P.n_slots + 8 * 4 * (int64_t)P.nblk +
         16 * 8 * (int64_t)P.vm_depth + 16 * 128 + 16 * cells + 4 * ((cells + 3) & ~3) + 16 +
         8 * (int64_t)(P.n + 2);
}

// Must match carve_pair() in lb_kernels.hip.
int64_t lds_bytes_pair(const DevPlan& P) {
  const int64_t g = P.g_rows;  // one block: its rows
  return g * 64 * 16 + 2 * 8 * (int64_t)(P.K + 1) * 16 + 2 * 8 * (int64_t)P.n_slots * 16 +
         2 * 8 * (int64_t)P.vm_depth * 16 + 64 * 16 + 128 * 16 + 8 * (int64_t)((P.n + 1) & ~1) + g * 64 * 4 + 16;
}

int set_lds_limit(const void* fn, size_t lds) {
  if (lds > 65536)
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return LB_OK;
}

}  // namespace

extern "C" const char* lb_last_error(void) { return g_error.c_str(); }

namespace {

// Closure of an initial support under the RHS program: an upper element can
// become nonzero only if it is nonzero at t0 or one of its terms reads an
// element that can.  Elements outside the closure stay exactly zero for the
// whole integration (every term that writes them reads zeros), so dropping
// them changes no value that is kept.
std::vector<uint8_t> active_closure(const PlanSource& S, std::vector<uint8_t> act) {
  for (bool grew = true; grew;) {
    grew = false;
    for (int o = 0; o < S.U; ++o) {
      if (act[o]) continue;
      for (int64_t k = S.output_ptrs[o]; k < S.output_ptrs[o + 1]; ++k)
        if (act[S.input_indices[k]]) { act[o] = 1; grew = true; break; }
    }
  }
  return act;
}

// Lane layout and device tables for the active elements `act` (lb_device.h).
int build_image(const PlanSource& src, const std::vector<uint8_t>& act, DevImage* im) {
  const int n = src.n, U = src.U;
  auto is_dyn = [&](int64_t k) { return src.coeff_indices[k] >= 0 && src.coef_dyn[src.coeff_indices[k]]; };
  // ---- split terms of active outputs into self (own element, time-independent
  //      coefficient) and gathers
  std::vector<int> n_gather(U, 0), n_self(U, 0);
  std::vector<uint8_t> published(U, 0);
  int64_t flops = 0;
  int n_act = 0;
  for (int o = 0; o < U; ++o) {
    if (!act[o]) continue;
    ++n_act;
    for (int64_t k = src.output_ptrs[o]; k < src.output_ptrs[o + 1]; ++k) {
      const int in = (int)src.input_indices[k];
      if (!act[in]) return fail(LB_ERR_INVALID, "active set is not closed under the RHS program");
      flops += (src.up_im[in] < 0) ? 4 : 8;
      if (is_dyn(k)) flops += 6;
      if (in == o && !is_dyn(k) && !src.input_conj[k]) {
        ++n_self[o];
      } else {
        ++n_gather[o];
        published[in] = 1;
      }
    }
  }
  // slots: active outputs ordered by gather count (descending), then index
  std::vector<int> order;
  for (int o = 0; o < U; ++o)
    if (act[o]) order.push_back(o);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n_gather[a] > n_gather[b]; });
  const int S = std::max(64, (n_act + 63) / 64 * 64), nblk = S / 64;
  // Blocks of 64 consecutive elements of that order, placed so that the
  // default launch's waves get about equal gather rows: wave w of nw holds
  // blocks w, w + nw, w + 2 nw, ..., so the blocks go, longest first, to the
  // wave with the fewest rows so far (at its next free position).  A block's
  // elements and their terms are unchanged; only its position moves.
  std::vector<int> slot_el(S, -1);  // slot -> element, or -1
  {
    std::vector<int> brows(nblk, 0);
    int rows_sum = 0;
    for (int b = 0; b < nblk; ++b) {
      for (int l = 0; l < 64 && 64 * b + l < n_act; ++l) brows[b] = std::max(brows[b], n_gather[order[64 * b + l]]);
      rows_sum += brows[b];
    }
    std::vector<int> pos_of(nblk);
    std::iota(pos_of.begin(), pos_of.end(), 0);
    int id = -1, W = 0, R = 1, maxw = 0, waves = 0;
    pick_config(S, rows_sum, &id, &W);
    if (nblk > 1 && id >= 0 && lb_dev_variant(id, &R, &maxw, &waves) && !lb_dev_variant_rows(id) && R > 1) {
      const int nw = W / 64;
      std::vector<int> load(nw, 0), next_r(nw, 0), cap(nw, 0);
      for (int w = 0; w < nw; ++w)
        while (w + nw * cap[w] < nblk) ++cap[w];
      for (int b = 0; b < nblk; ++b) {  // brows is non-increasing in b
        // the least loaded wave with a free position; of equal loads, the
        // one with the fewest free positions (it can take the fewest blocks)
        int best = -1;
        for (int w = 0; w < nw; ++w) {
          if (next_r[w] >= cap[w]) continue;
          if (best < 0 || load[w] < load[best] ||
              (load[w] == load[best] && cap[w] - next_r[w] < cap[best] - next_r[best]))
            best = w;
        }
        pos_of[b] = best + nw * next_r[best]++;
        load[best] += brows[b];
      }
    }
    for (int s0 = 0; s0 < n_act; ++s0) slot_el[64 * pos_of[s0 / 64] + s0 % 64] = order[s0];
  }
  std::vector<int32_t> slot_up(S, -1), slot_re(S, -1), slot_im(S, -1), slot_pub(S, -1);
  std::vector<int32_t> pub_of_upper(U, -1);
  int n_pub = 0;
  // A plan of one block publishes every element at its slot index, so the
  // gather words name the owning lane (the single-wave kernels read inputs
  // by cross-lane permutes).
  const bool one_block = S == 64;
  for (int s = 0; s < S; ++s) {
    const int o = slot_el[s];
    if (o < 0) continue;
    slot_up[s] = o;
    slot_re[s] = src.up_re[o];
    slot_im[s] = src.up_im[o];
    if (one_block) { pub_of_upper[o] = s; slot_pub[s] = s; n_pub = s + 1; }
    else if (published[o]) { pub_of_upper[o] = n_pub; slot_pub[s] = n_pub++; }
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
      const int o = slot_el[64 * b + l];
      if (o >= 0) rows = std::max(rows, n_gather[o]);
    }
    blk_row0[b] = rows_total;
    blk_rows[b] = rows;
    rows_total += rows;
  }
  std::vector<uint32_t> g_word((size_t)rows_total * 64, lb_gword(0, true, false, false, LB_STATIC_COEF));
  std::vector<double> g_f((size_t)rows_total * 64 * 2, 0.0);
  for (int s = 0; s < S; ++s) {
    const int o = slot_el[s], b = s / 64, l = s % 64;
    if (o < 0) continue;
    int js = 0, jg = 0;
    for (int64_t k = src.output_ptrs[o]; k < src.output_ptrs[o + 1]; ++k) {
      const int in = (int)src.input_indices[k];
      const int64_t ci = src.coeff_indices[k];
      if (in == o && !is_dyn(k) && !src.input_conj[k]) {
        const size_t e = (size_t)s * max_self + js++;
        self_coef[e] = (int32_t)ci;
        self_conj[e] = src.coeff_conj[k];
        self_f[2 * e] = src.factors_re[k];
        self_f[2 * e + 1] = src.factors_im[k];
      } else {
        const size_t e = (size_t)(blk_row0[b] + jg++) * 64 + l;
        g_word[e] = lb_gword((uint32_t)pub_of_upper[in], src.up_im[in] < 0, src.input_conj[k] != 0,
                             src.coeff_conj[k] != 0, ci < 0 ? LB_STATIC_COEF : (uint32_t)ci);
        g_f[2 * e] = src.factors_re[k];
        g_f[2 * e + 1] = src.factors_im[k];
      }
    }
  }
  // transposed tables of a differentiable plan: for each input slot, the terms
  // that read it (self-folded terms excluded), output named by its slot
  std::vector<int32_t> t_row0(nblk, 0), t_rows(nblk, 0);
  std::vector<uint32_t> t_word;
  std::vector<double> t_f;
  int t_total = 0;
  if (src.n_diff > 0) {
    std::vector<int32_t> slot_of(U, -1);
    for (int s = 0; s < S; ++s)
      if (slot_el[s] >= 0) slot_of[slot_el[s]] = s;
    std::vector<std::vector<int64_t>> readers(S);
    for (int o = 0; o < U; ++o) {
      if (!act[o]) continue;
      for (int64_t k = src.output_ptrs[o]; k < src.output_ptrs[o + 1]; ++k) {
        const int in = (int)src.input_indices[k];
        if (in == o && !is_dyn(k) && !src.input_conj[k]) continue;
        readers[slot_of[in]].push_back(k);
      }
    }
    for (int b = 0; b < nblk; ++b) {
      int rows = 0;
      for (int l = 0; l < 64; ++l) rows = std::max<int>(rows, (int)readers[64 * b + l].size());
      t_row0[b] = t_total;
      t_rows[b] = rows;
      t_total += rows;
    }
    t_word.assign((size_t)t_total * 64, lb_gword(0, true, false, false, LB_STATIC_COEF));
    t_f.assign((size_t)t_total * 64 * 2, 0.0);
    std::vector<int32_t> out_of(src.T, -1);
    for (int o = 0; o < U; ++o)
      for (int64_t k = src.output_ptrs[o]; k < src.output_ptrs[o + 1]; ++k) out_of[k] = o;
    for (int s = 0; s < S; ++s) {
      const int b = s / 64, l = s % 64, in = slot_el[s];
      if (in < 0) continue;
      int j = 0;
      for (int64_t k : readers[s]) {
        const int64_t ci = src.coeff_indices[k];
        const size_t e = (size_t)(t_row0[b] + j++) * 64 + l;
        t_word[e] = lb_gword((uint32_t)slot_of[out_of[k]], src.up_im[in] < 0, src.input_conj[k] != 0,
                             src.coeff_conj[k] != 0, ci < 0 ? LB_STATIC_COEF : (uint32_t)ci);
        t_f[2 * e] = src.factors_re[k];
        t_f[2 * e + 1] = src.factors_im[k];
      }
    }
  }
  // packed index -> upper index, and the active flags (outputs of dropped
  // elements are written as zeros)
  std::vector<int32_t> packed_upper((size_t)n * n, 0);
  for (int u = 0; u < U; ++u) {
    packed_upper[src.up_re[u]] = u;
    if (src.up_im[u] >= 0) packed_upper[src.up_im[u]] = u;
  }

  DevPlan& P = im->dev;
  P.n = n;
  P.dim = n * n;
  P.upper_len = U;
  P.S = S;
  P.nblk = nblk;
  P.n_pub = n_pub;
  P.K = src.K;
  P.KD = src.K * (1 + src.n_diff);
  P.n_diff = src.n_diff;
  P.diff_time = src.diff_time;
  P.n_slots = src.prog.n_slots;
  P.n_base = src.n_base;
  P.n_compounds = src.prog.n_compounds;
  P.vm_depth = std::max(src.prog.max_depth, 1);
  P.g_rows = rows_total;
  P.lds_tables = (int64_t)rows_total * 64 * 20 <= kLdsTableBytes ? 1 : 0;
  P.max_self = max_self;
  P.time_dependent = src.time_dependent ? 1 : 0;
  P.restricted = n_act < U ? 1 : 0;
  int rc = LB_OK;
  auto up = [&](auto& v, auto** dst) {
    if (rc == LB_OK) rc = upload(im, v, dst);
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
  {
    std::vector<DevOp> ops(src.prog.instr.size());
    for (size_t i = 0; i < ops.size(); ++i) {
      const DevInstr& q = src.prog.instr[i];
      ops[i].code = (q.op & 0xFF) | ((q.argc & 0xFF) << 8) | ((q.function & 0xFFFF) << 16);
      ops[i].index = q.index;
      ops[i].re = q.re;
      ops[i].im = q.im;
      ops[i].aux = q.aux;
    }
    up(ops, &P.instr);
  }
  up(src.prog.expr_start, &P.expr_start);
  up(src.compound_slot, &P.compound_slot);
  up(src.expr_dyn, &P.expr_dyn);
  {
    // Evaluation order: the compounds the solve needs, the K coefficients,
    // then the compounds only the derivative programs read and the
    // derivative programs.  Each part keeps the program's order, so every
    // expression follows the compounds it reads; a solve evaluates the first
    // two parts only.
    const int nc = src.prog.n_compounds, K = src.K, KD = src.K * (1 + src.n_diff);
    std::vector<int> order;
    for (int e = 0; e < nc; ++e)
      if (src.fwd_compound[e]) order.push_back(e);
    for (int e = nc; e < nc + K; ++e) order.push_back(e);
    const int n_fwd = (int)order.size();
    for (int e = 0; e < nc; ++e)
      if (!src.fwd_compound[e]) order.push_back(e);
    for (int e = nc + K; e < nc + KD; ++e) order.push_back(e);
    std::vector<DevExpr> all, dyn;
    int n_dyn_fwd = 0;
    for (int i = 0; i < (int)order.size(); ++i) {
      const int e = order[i];
      DevExpr r{};
      r.start = src.prog.expr_start[e];
      r.end = src.prog.expr_start[e + 1];
      if (e < nc) {
        r.dst = src.prog.compound_slot[e];
      } else {
        const int j = e - nc;
        r.dst = -((j < K ? j : j + 1) + 1);
      }
      all.push_back(r);
      if (src.expr_dyn[e]) {
        dyn.push_back(r);
        if (i < n_fwd) ++n_dyn_fwd;
      }
    }
    P.n_all_fwd = n_fwd;
    P.n_all = nc + KD;
    P.n_dyn_fwd = n_dyn_fwd;
    P.n_dyn = (int32_t)dyn.size();
    if (all.empty()) all.push_back(DevExpr{});
    if (dyn.empty()) dyn.push_back(DevExpr{});
    up(all, &P.ex_all);
    up(dyn, &P.ex_dyn);
  }
  up(src.base, &P.base);
  up(act, &P.upper_active);
  up(packed_upper, &P.packed_upper);
  P.t_rows_total = t_total;
  up(t_row0, &P.t_blk_row0);
  up(t_rows, &P.t_blk_rows);
  up(t_word, &P.t_word);
  up(t_f, &P.t_f);
  if (rc != LB_OK) return rc;
  pick_config(S, rows_total, &im->default_variant, &im->default_W);
  // paired variant: the fewest register rows holding a one-block plan of <= 32 elements
  im->pair_variant = -1;
  if (nblk == 1 && n_act <= 32) {
    int best_g = 1 << 30;
    for (int id = 0; id < lb_dev_variant_count(); ++id) {
      const int g = lb_dev_variant_rows(id);
      if (lb_dev_variant_pair(id) && rows_total <= g && g < best_g) { best_g = g; im->pair_variant = id; }
    }
  }
  im->lds_bytes = lds_bytes_for(P);
  im->n_pub = n_pub;
  im->rows_total = rows_total;
  im->g_rows_host = blk_rows;
  im->t_rows_host = t_rows.empty() ? 0 : t_rows[0];
  im->flops = flops;
  im->active = n_act;
  im->active_mask = act;
  for (int o = 0; o < U; ++o) { im->self_terms += n_self[o]; im->gather_terms += n_gather[o]; }
  return LB_OK;
}

}  // namespace

extern "C" int lb_plan_create(const lb_plan_desc* d, lb_plan** out_plan) {
  if (!out_plan) return fail(LB_ERR_INVALID, "out_plan is null");
  *out_plan = nullptr;
  if (!d || d->n_states <= 0) return fail(LB_ERR_INVALID, "n_states must be positive");
  const int n = d->n_states, U = n * (n + 1) / 2;
  if (d->n_base < 0 || d->n_base > d->n_slots) return fail(LB_ERR_INVALID, "n_base exceeds n_slots");
  if (d->n_coefficients > 254) return fail(LB_ERR_UNSUPPORTED, "more than 254 Hamiltonian coefficients");
  if (d->diff_time != 0 && d->diff_time != 1) return fail(LB_ERR_INVALID, "diff_time must be 0 or 1");
  if (d->diff_time && d->n_diff_params <= 0)
    return fail(LB_ERR_INVALID, "a time derivative column needs differentiated parameters");
  if (d->n_diff_params < 0 || d->n_diff_params + d->diff_time > LB_MAX_DIFF)
    return fail(LB_ERR_UNSUPPORTED, "at most 8 derivative columns (parameters plus time)");
  for (int m = 0; m < d->n_diff_params; ++m)
    if (!d->diff_param_slots || d->diff_param_slots[m] < 0 || d->diff_param_slots[m] >= d->n_base)
      return fail(LB_ERR_INVALID, "differentiated parameters must be base slots");
  // the expression list seen by the program builder: compounds, coefficients,
  // then the derivative programs of a differentiable plan
  lb_plan_desc dd = *d;
  dd.n_coefficients = d->n_coefficients * (1 + d->n_diff_params + d->diff_time);
  const int n_expr = d->n_compounds + dd.n_coefficients;
  for (int c = 0; c < d->n_compounds; ++c)
    if (d->compound_slots[c] < d->n_base || d->compound_slots[c] >= d->n_slots)
      return fail(LB_ERR_INVALID, "compound slot out of range");
  for (int e = 0; e < n_expr; ++e) {
    std::string why;
    int rc = check_expression(&dd, dd.expr_ptrs[e], dd.expr_ptrs[e + 1], dd.n_slots, why);
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

  std::unique_ptr<lb_plan> p(new lb_plan());
  PlanSource& src = p->src;
  src.n = n;
  src.U = U;
  src.K = d->n_coefficients;
  src.n_diff = d->n_diff_params + d->diff_time;
  src.diff_time = d->diff_time;
  src.n_base = d->n_base;
  src.n_slots = d->n_slots;
  src.T = T;
  src.output_ptrs.assign(d->output_ptrs, d->output_ptrs + U + 1);
  src.input_indices.assign(d->input_indices, d->input_indices + T);
  src.coeff_indices.assign(d->coeff_indices, d->coeff_indices + T);
  src.coeff_conj.assign(d->coeff_conj, d->coeff_conj + T);
  src.input_conj.assign(d->input_conj, d->input_conj + T);
  src.factors_re.assign(d->factors_re, d->factors_re + T);
  src.factors_im.assign(d->factors_im, d->factors_im + T);

  // ---- time dependence (plan.rs:1235-1291)
  std::vector<uint8_t> dyn_slot(d->n_slots, 0), compound_dyn(d->n_compounds, 0), coef_dyn(dd.n_coefficients, 0);
  bool time_dependent = false;
  auto reads_dynamic = [&](int e, bool* uses_time) {
    bool dyn = false;
    *uses_time = false;
    for (int q = d->expr_ptrs[e]; q < d->expr_ptrs[e + 1]; ++q) {
      if (d->instr_op[q] == 4) { dyn = true; *uses_time = true; }
      if (d->instr_op[q] == 2 && dyn_slot[d->instr_index[q]]) dyn = true;
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
  for (int k = 0; k < dd.n_coefficients; ++k) {
    bool ut;
    coef_dyn[k] = reads_dynamic(d->n_compounds + k, &ut);
    time_dependent |= ut;
    if (k < d->n_coefficients) n_dyn_coef += coef_dyn[k];
  }
  src.coef_dyn = coef_dyn;
  src.time_dependent = time_dependent;
  src.n_dyn_coef = n_dyn_coef;

  // ---- upper index -> packed indices (rhs.rs:151-176)
  src.up_re.assign(U, 0);
  src.up_im.assign(U, 0);
  {
    int cursor = n, u = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++u) {
        if (i == j) { src.up_re[u] = i; src.up_im[u] = -1; }
        else { src.up_re[u] = cursor; src.up_im[u] = cursor + 1; cursor += 2; }
      }
  }

  // ---- expression machine (time-independent subexpressions cached per trajectory)
  {
    std::string why;
    if (lb_build_program(&dd, dyn_slot, compound_dyn, coef_dyn, &src.prog, why) != 0)
      return fail(LB_ERR_INVALID, why);
  }
  src.expr_dyn.assign(src.prog.compound_dyn.begin(), src.prog.compound_dyn.end());
  src.expr_dyn.insert(src.expr_dyn.end(), src.prog.coef_dyn.begin(), src.prog.coef_dyn.end());
  {
    // The compounds a solve leaves out: those only the derivative programs of
    // a differentiable plan read (directly or through other compounds).
    // Every other compound -- read by a coefficient value, or by nothing the
    // plan holds (a stop-event expression may read it) -- is evaluated by the
    // solve as before.
    const int nc = src.prog.n_compounds, KD = src.K * (1 + src.n_diff);
    std::vector<int> comp_of_slot(std::max(src.prog.n_slots, 1), -1);
    for (int c = 0; c < nc; ++c) comp_of_slot[src.prog.compound_slot[c]] = c;
    // the compounds read, directly or through other compounds, by expressions `from`
    auto reach = [&](const std::vector<int>& from) {
      std::vector<uint8_t> seen(nc, 0);
      std::vector<int> todo = from;
      while (!todo.empty()) {
        const int e = todo.back();
        todo.pop_back();
        for (int q = src.prog.expr_start[e]; q < src.prog.expr_start[e + 1]; ++q) {
          const DevInstr& ins = src.prog.instr[q];
          if (ins.op != 2 || ins.index < 0 || ins.index >= (int)comp_of_slot.size()) continue;
          const int c = comp_of_slot[ins.index];
          if (c >= 0 && !seen[c]) {
            seen[c] = 1;
            todo.push_back(c);
          }
        }
      }
      return seen;
    };
    std::vector<int> values, derivs;
    for (int e = nc; e < nc + src.K; ++e) values.push_back(e);
    for (int e = nc + src.K; e < nc + KD; ++e) derivs.push_back(e);
    const std::vector<uint8_t> by_values = reach(values), by_derivs = reach(derivs);
    src.fwd_compound.assign(nc, 1);
    for (int c = 0; c < nc; ++c)
      if (by_derivs[c] && !by_values[c]) src.fwd_compound[c] = 0;
    // what a kept compound reads is kept too
    std::vector<int> kept;
    for (int c = 0; c < nc; ++c)
      if (src.fwd_compound[c]) kept.push_back(c);
    const std::vector<uint8_t> by_kept = reach(kept);
    for (int c = 0; c < nc; ++c)
      if (by_kept[c]) src.fwd_compound[c] = 1;
  }
  if (src.expr_dyn.empty()) src.expr_dyn.push_back(0);
  src.compound_slot = src.prog.compound_slot;
  if (src.compound_slot.empty()) src.compound_slot.push_back(0);
  src.base.assign(2 * (size_t)d->n_base, 0.0);
  for (int i = 0; i < d->n_base; ++i) { src.base[2 * i] = d->base_re[i]; src.base[2 * i + 1] = d->base_im[i]; }

  // ---- device side
  HIP_TRY(hipGetDevice(&p->device));
  HIP_TRY(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device));
  if (int rc0 = p->queues.init()) return rc0;
  p->full.reset(new DevImage());
  int rc = build_image(src, std::vector<uint8_t>(U, 1), p->full.get());
  if (rc != LB_OK) return rc;
  const DevImage& F = *p->full;
  int64_t* info = p->info;
  info[0] = n; info[1] = (int64_t)n * n; info[2] = U; info[3] = T; info[4] = d->n_coefficients;
  info[5] = time_dependent; info[6] = F.n_pub; info[7] = F.gather_terms; info[8] = F.self_terms;
  info[9] = F.dev.S; info[10] = F.rows_total; info[11] = F.flops; info[12] = n_dyn_coef; info[13] = F.lds_bytes;
  info[14] = F.default_variant + 1; info[15] = F.default_W;
  *out_plan = p.release();
  return LB_OK;
}

extern "C" void lb_plan_destroy(lb_plan* p) { delete p; }

extern "C" int lb_plan_specialize(lb_plan* p, int32_t enable) {
  if (!p) return fail(LB_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lock(p->mu);
  if (!enable || !p->src.time_dependent) {  // nothing is evaluated per step without time dependence
    p->spec = false;
    return LB_OK;
  }
  if (p->spec_source.empty()) {
    std::string why;
    // only the programs a solve evaluates (the derivative-only compounds of a
    // differentiable plan stay out)
    std::vector<int32_t> dyn = p->src.expr_dyn;
    for (int c = 0; c < p->src.prog.n_compounds; ++c)
      if (!p->src.fwd_compound[c]) dyn[c] = 0;
    p->spec_source = lb_rtc_coef_source(p->src.prog, dyn, p->src.K, why);
    if (p->spec_source.empty()) return fail(LB_ERR_UNSUPPORTED, why);
  }
  p->spec = true;
  return LB_OK;
}

extern "C" int lb_plan_specialize_status(const lb_plan* p, char* buf, int64_t cap) {
  if (!p) return fail(LB_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lock(const_cast<lb_plan*>(p)->mu);
  if (p->spec_failure.empty()) return 0;
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>((size_t)cap - 1, p->spec_failure.size());
    std::memcpy(buf, p->spec_failure.data(), n);
    buf[n] = '\0';
  }
  return 1;
}

extern "C" int lb_plan_info(const lb_plan* p, int64_t* info) {
  if (!p || !info) return fail(LB_ERR_INVALID, "null argument");
  std::memcpy(info, p->info, sizeof(p->info));
  return LB_OK;
}

namespace {
int select_image(lb_plan* p, const uint8_t* y0_support, const DevImage** out, bool may_build = true);
}

extern "C" int lb_plan_support_info(const lb_plan* p, const uint8_t* y0_support, int64_t* info) {
  if (!p || !info) return fail(LB_ERR_INVALID, "null argument");
  const DevImage* im = nullptr;
  int rc = select_image(const_cast<lb_plan*>(p), y0_support, &im);
  if (rc) return rc;
  std::memcpy(info, p->info, sizeof(p->info));
  int64_t terms = 0, packed = 0;  // terms and packed reals of the active outputs
  for (int o = 0; o < p->src.U; ++o)
    if (im->active_mask[o]) {
      terms += p->src.output_ptrs[o + 1] - p->src.output_ptrs[o];
      packed += p->src.up_im[o] < 0 ? 1 : 2;
    }
  info[1] = packed;
  info[3] = terms;
  info[2] = im->active;
  info[6] = im->n_pub; info[7] = im->gather_terms; info[8] = im->self_terms;
  info[9] = im->dev.S; info[10] = im->rows_total; info[11] = im->flops; info[13] = im->lds_bytes;
  info[14] = im->default_variant + 1; info[15] = im->default_W;
  return LB_OK;
}

namespace {

// variant: 0 = plan default, k > 0 = kernel variant k - 1.  threads: 0 = the
// variant's smallest covering workgroup.
// pair_ok: the call may run two trajectories per wave (persistent, pushes at
// every accepted step, dt given, no trace / save plan / tape / replay).
int resolve_config(const DevImage* p, int variant, int threads, bool pair_ok, int* id, int* W) {
  if (threads < 0 || threads % 64 != 0)
    return fail(LB_ERR_INVALID, "threads_per_traj must be a multiple of 64");
  if (variant <= 0 && threads == 0 && pair_ok && p->pair_variant >= 0) {
    *id = p->pair_variant;
    *W = 64;
    return LB_OK;
  }
  if (variant > 0 && lb_dev_variant_pair(variant - 1)) {
    const int g = lb_dev_variant_rows(variant - 1);
    if (!pair_ok || p->pair_variant < 0 || p->rows_total > g)
      return fail(LB_ERR_UNSUPPORTED,
                  "a paired variant needs a plan of at most 32 elements (and its gather rows) and a persistent "
                  "call pushing at every accepted step with dt given (no trace, save plan, tape or replay)");
    *id = variant - 1;
    *W = 64;
    return LB_OK;
  }
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
    *W = threads ? threads : variant_threads(*id, p->dev.S, p->dev.g_rows);
    if (lb_dev_variant_rows(*id) && (*W != 64 || !variant_threads(*id, p->dev.S, p->dev.g_rows)))
      return fail(LB_ERR_UNSUPPORTED, "single-wave variant needs a plan of at most 64 elements");
    if (!*W || *W > maxw || (int64_t)R * *W < p->dev.S)
      return fail(LB_ERR_UNSUPPORTED, "kernel variant does not cover the plan with this workgroup size");
    return LB_OK;
  }
  // threads given: the variant with the fewest slots per thread that covers the plan
  for (int v = 0; v < lb_dev_variant_count(); ++v) {
    lb_dev_variant(v, &R, &maxw, &waves);
    if (lb_dev_variant_rows(v)) continue;
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
    if (slots[i] < 0 || slots[i] >= p->src.n_base)
      return fail(LB_ERR_INVALID, "parameter override slot " + std::to_string(slots[i]) +
                                      " is not a base parameter slot");
    dst[i] = slots[i];
  }
  return LB_OK;
}

// The device image for a solve: the full plan, or (with an initial-support
// mask over packed indices) the plan restricted to that support's closure,
// built on first use (unless `may_build` is false) and cached.
int select_image(lb_plan* p, const uint8_t* y0_support, const DevImage** out, bool may_build) {
  *out = p->full.get();
  if (!y0_support) return LB_OK;
  const PlanSource& S = p->src;
  std::vector<uint8_t> seed(S.U, 0);
  for (int u = 0; u < S.U; ++u)
    seed[u] = (y0_support[S.up_re[u]] || (S.up_im[u] >= 0 && y0_support[S.up_im[u]])) ? 1 : 0;
  std::vector<uint8_t> act = active_closure(S, seed);
  if (std::all_of(act.begin(), act.end(), [](uint8_t v) { return v != 0; })) return LB_OK;
  std::lock_guard<std::mutex> lock(p->mu);
  auto it = p->restricted.find(act);
  if (it != p->restricted.end()) {
    *out = it->second.get();
    return LB_OK;
  }
  if (!may_build)
    return fail(LB_ERR_INVALID,
                "no device tables for this y0 support yet (an lb_solve with the same support builds them)");
  std::unique_ptr<DevImage> im(new DevImage());
  int rc = build_image(S, act, im.get());
  if (rc != LB_OK) {
    for (void* a : im->allocations) (void)hipFree(a);
    return rc;
  }
  *out = im.get();
  p->restricted.emplace(std::move(act), std::move(im));
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
  const DevImage* im = p->full.get();
  if ((rc = resolve_config(im, 0, 0, false, &id, &W))) return rc;
  void* fn = lb_dev_rhs_kernel(id);
  DevPlan P = im->dev;
  void* args[] = {&P, &c};
  if ((rc = set_lds_limit(fn, (size_t)im->lds_bytes))) return rc;
  int grid = (int)std::min<int64_t>(batch, 4096);
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, (size_t)im->lds_bytes, (hipStream_t)stream));
  HIP_TRY(hipGetLastError());
  return LB_OK;
}

namespace {

// Validate a solve request and fill the kernel's call record; picks the
// device image (full or restricted to the y0 support closure).
int fill_call(const lb_plan* p, const lb_solve_args* a, const DevImage** im_out, DevCall* cp,
              bool may_build = true) {
  if (!a->y0 || !a->out_values || !a->out_stats || !a->out_fail_t)
    return fail(LB_ERR_INVALID, "y0 and output buffers are required");
  if (a->maxiters <= 0) return fail(LB_ERR_INVALID, "maxiters must be positive");
  if (a->solver < LB_SOLVER_DOPRI5 || a->solver > LB_SOLVER_FIXED_RK4)
    return fail(LB_ERR_INVALID, "solver must be dopri5, tsit5, fixed_dopri5, fixed_rk2 or fixed_rk4");
  if (!std::isfinite(a->t0) || (!a->t1_per_traj && !std::isfinite(a->t1)))
    return fail(LB_ERR_INVALID, "t_span must be finite");
  if (a->n_traj > INT32_MAX) return fail(LB_ERR_UNSUPPORTED, "at most 2^31-1 trajectories per call");
  const DevImage* im = nullptr;
  int rc = select_image(const_cast<lb_plan*>(p), a->y0_support, &im, may_build);
  if (rc) return rc;
  DevCall& c = *cp;
  c = DevCall{};
  c.n_traj = a->n_traj;
  c.t0 = a->t0;
  c.t1 = a->t1;
  c.t1_per_traj = a->t1_per_traj;
  c.n_total = a->n_traj;
  c.queue_order = a->queue_order;
  c.abstol = a->abstol;
  c.reltol = a->reltol;
  c.dt = a->dt;
  c.maxiters = a->maxiters;
  c.save_start = a->save_start;
  c.solver = a->solver;
  c.y0 = a->y0;
  c.y0_shared = a->y0_shared;
  if ((rc = fill_overrides(p, a->n_override, a->override_slots, c.override_slot))) return rc;
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
  if (a->output_kind < 0 || a->output_kind > 4) return fail(LB_ERR_INVALID, "unknown output kind");
  if (a->output_kind == LB_OUT_RATE && (!a->trace || a->n_weights <= 0))
    return fail(LB_ERR_INVALID, "rate output needs trace mode and weights");
  if (a->n_weights > LB_MAX_WEIGHTS) return fail(LB_ERR_UNSUPPORTED, "at most 64 integral weights");
  c.n_weights = a->n_weights;
  for (int i = 0; i < a->n_weights; ++i) {
    if (a->weight_indices[i] < 0 || a->weight_indices[i] >= p->src.n)
      return fail(LB_ERR_INVALID, "integral weight index out of range");
    c.weight_index[i] = a->weight_indices[i];
    c.weight_value[i] = a->weight_values[i];
  }
  if (a->output_kind == LB_OUT_SELECTED) {
    if (a->n_selected > LB_MAX_SELECTED) return fail(LB_ERR_UNSUPPORTED, "at most 128 selected entries");
    const int n = p->src.n;
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
  if (a->n_saveat < 0) return fail(LB_ERR_INVALID, "n_saveat must be non-negative");
  if (a->n_saveat > 0 && !a->saveat) return fail(LB_ERR_INVALID, "saveat times are required");
  c.n_saveat = a->n_saveat;
  c.saveat = a->saveat;
  c.trace = a->trace ? 1 : 0;
  if (c.trace) {
    if (a->capacity <= 0 || !a->out_counts)
      return fail(LB_ERR_INVALID, "trace output needs a positive capacity and out_counts");
    c.capacity = a->capacity;
    c.out_times = a->out_times;
    c.out_counts = a->out_counts;
  }
  if (a->tape_h) {
    if (!a->tape_t || !a->tape_c || !a->tape_f || !a->tape_y || !a->tape_n || a->tape_cap <= 0 ||
        a->ckpt_every <= 0)
      return fail(LB_ERR_INVALID,
                  "an adjoint tape needs tape_t, tape_c, tape_f, tape_y, tape_n, tape_cap and ckpt_every");
    c.tape_t = a->tape_t;
    c.tape_h = a->tape_h;
    c.tape_c = a->tape_c;
    c.tape_f = a->tape_f;
    c.tape_y = a->tape_y;
    c.tape_n = a->tape_n;
    c.tape_cap = a->tape_cap;
    c.ckpt_every = a->ckpt_every;
  }
  if (a->replay_h) {
    if (!a->replay_n || a->tape_cap <= 0) return fail(LB_ERR_INVALID, "a replay needs replay_n and tape_cap");
    c.replay_h = a->replay_h;
    c.replay_n = a->replay_n;
    c.tape_cap = a->tape_cap;
  }
  c.out_event_t = a->out_event_t;
  *im_out = im;
  return LB_OK;
}

// The stop event of a call: checks it and copies its program (population:
// the diagonal indices; expression: the instructions as device records) into
// the plan's event buffer on `stream`, through pinned staging whose previous
// copy is waited for first.  *depth: the expression's stack depth.
int prepare_event(lb_plan* p, const lb_solve_args* a, hipStream_t stream, DevCall& c, int* depth) {
  *depth = 0;
  if (a->event_kind == LB_EVENT_NONE) return LB_OK;
  if (a->event_kind != LB_EVENT_POPULATION && a->event_kind != LB_EVENT_EXPRESSION)
    return fail(LB_ERR_INVALID, "unknown stop event kind " + std::to_string(a->event_kind));
  if (a->tape_h || a->replay_h)
    return fail(LB_ERR_UNSUPPORTED, "stop events are not supported with adjoint tapes or replays");
  std::vector<unsigned char> bytes;
  int n = 0;
  if (a->event_kind == LB_EVENT_POPULATION) {
    n = a->n_event_indices;
    if (n <= 0 || !a->event_indices) return fail(LB_ERR_INVALID, "population event requires at least one index");
    for (int k = 0; k < n; ++k)
      if (a->event_indices[k] < 0 || a->event_indices[k] >= p->src.n)
        return fail(LB_ERR_INVALID, "population event index " + std::to_string(a->event_indices[k]) +
                                        " out of bounds");
    const unsigned char* b = reinterpret_cast<const unsigned char*>(a->event_indices);
    bytes.assign(b, b + (size_t)n * sizeof(int32_t));
  } else {
    n = a->n_event_instr;
    if (n <= 0 || !a->event_op || !a->event_index || !a->event_argc || !a->event_function || !a->event_re ||
        !a->event_im)
      return fail(LB_ERR_INVALID, "the stop event expression is empty");
    lb_plan_desc d{};
    d.instr_op = a->event_op;
    d.instr_index = a->event_index;
    d.instr_argc = a->event_argc;
    d.instr_function = a->event_function;
    d.instr_re = a->event_re;
    d.instr_im = a->event_im;
    std::string why;
    const int rc = check_expression(&d, 0, n, p->src.n_slots, why, depth);
    if (rc == -1) return fail(LB_ERR_INVALID, "stop event expression: " + why);
    if (rc == -2) return fail(LB_ERR_UNSUPPORTED, "stop event expression: " + why);
    std::vector<DevOp> ops(n);
    for (int i = 0; i < n; ++i) {
      ops[i].code = (a->event_op[i] & 0xFF) | ((a->event_argc[i] & 0xFF) << 8) | ((a->event_function[i] & 0xFFFF) << 16);
      ops[i].index = a->event_index[i];
      ops[i].re = a->event_re[i];
      ops[i].im = a->event_im[i];
      ops[i].aux = 0.0;
    }
    const unsigned char* b = reinterpret_cast<const unsigned char*>(ops.data());
    bytes.assign(b, b + ops.size() * sizeof(DevOp));
  }
  std::lock_guard<std::mutex> lock(p->mu);
  if (!p->ev_copied) HIP_TRY(hipEventCreateWithFlags(&p->ev_copied, hipEventDisableTiming));
  else HIP_TRY(hipEventSynchronize(p->ev_copied));
  if (bytes.size() > p->ev_cap) {
    // earlier calls on this stream may still read the old buffer
    HIP_TRY(hipStreamSynchronize(stream));
    if (p->ev_dev) HIP_TRY(hipFree(p->ev_dev));
    if (p->ev_pinned) HIP_TRY(hipHostFree(p->ev_pinned));
    p->ev_dev = p->ev_pinned = nullptr;
    p->ev_cap = 0;
    const size_t cap = std::max<size_t>(bytes.size(), 4096);
    HIP_TRY(hipMalloc(&p->ev_dev, cap));
    HIP_TRY(hipHostMalloc(&p->ev_pinned, cap, hipHostMallocDefault));
    p->ev_cap = cap;
  }
  std::memcpy(p->ev_pinned, bytes.data(), bytes.size());
  HIP_TRY(hipMemcpyAsync(p->ev_dev, p->ev_pinned, bytes.size(), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipEventRecord(p->ev_copied, stream));
  c.event_kind = a->event_kind;
  c.n_event = n;
  c.event_threshold = a->event_threshold;
  if (a->event_kind == LB_EVENT_POPULATION) c.event_idx = static_cast<const int32_t*>(p->ev_dev);
  else c.event_ops = static_cast<const DevOp*>(p->ev_dev);
  return LB_OK;
}

}  // namespace

namespace {

// Fixed-step solvers and calls with a stop event: the multi-wave kernel of
// that family with the fewest slots per thread that covers the plan (or the
// given threads per trajectory; `variant` is not used).
int launch_table(const lb_plan* p, const lb_solve_args* a, const DevImage* im, DevCall& c, void* stream,
                 bool event, int ev_depth) {
  if (a->tape_h || a->replay_h)
    return fail(LB_ERR_UNSUPPORTED, "adjoint tapes and replays need an adaptive solver (dopri5 or tsit5)");
  const int S = im->dev.S;
  int id = -1, W = 0, R = 0, maxw = 0;
  for (int v = 0; event ? lb_dev_event_variant(v, &R, &maxw) : lb_dev_fixed_variant(v, &R, &maxw); ++v) {
    const int w = a->threads_per_traj > 0 ? a->threads_per_traj : std::max(64, ((S + R - 1) / R + 63) / 64 * 64);
    if (w <= maxw && (int64_t)R * w >= S) { id = v; W = w; break; }
  }
  if (id < 0) return fail(LB_ERR_UNSUPPORTED, event ? "no stop-event kernel variant covers the plan"
                                                    : "no fixed-step kernel variant covers the plan");
  void* fn = event ? lb_dev_event_kernel(id, a->solver) : lb_dev_fixed_kernel(id, a->solver);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "kernel variant is not built");
  void* reset = lb_dev_queue_reset_kernel();
  hipStream_t s = (hipStream_t)stream;
  DevPlan P = im->dev;
  P.vm_depth = std::max(P.vm_depth, ev_depth);  // the stop-event expression runs on lane 0's stack
  const size_t lds = (size_t)lds_bytes_for(P);
  int rc = set_lds_limit(fn, lds);
  if (rc) return rc;
  int slot = 0;
  if ((rc = const_cast<lb_plan*>(p)->queues.acquire(s, &c.queue, &slot))) return rc;
  if (a->mode == LB_MODE_PLAIN) {
    for (int64_t i = 0; i < a->n_traj; ++i) {
      c.n_traj = 1;
      c.traj_offset = i;
      void* rargs[] = {&c.queue};
      HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
      void* args[] = {&P, &c};
      HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(W), args, lds, s));
    }
    HIP_TRY(hipGetLastError());
    return const_cast<lb_plan*>(p)->queues.retire(slot, s);
  }
  int grid = a->workgroups;
  if (grid <= 0) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
    grid = std::max(1, per_cu) * std::max(1, p->cus);
  }
  grid = (int)std::min<int64_t>(grid, a->n_traj);
  c.traj_offset = 0;
  void* rargs[] = {&c.queue};
  HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
  void* args[] = {&P, &c};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
  HIP_TRY(hipGetLastError());
  return const_cast<lb_plan*>(p)->queues.retire(slot, s);
}

}  // namespace

namespace {

// The plan-specialised form of paired variant `id` for `solver`, compiled on
// first use; nullptr when the plan is not specialised.  (Only the paired
// kernels take the generated code: a specialised multi-wave kernel launched
// with 1024-thread workgroups had its queue aborted at dispatch, cause not
// found, so the single- and multi-wave kernels stay on the precompiled form.)
int spec_kernel(lb_plan* p, int id, int solver, int rows, hipStream_t stream, hipFunction_t* fn) {
  *fn = nullptr;
  std::lock_guard<std::mutex> lock(p->mu);
  if (!p->spec) return LB_OK;
  // the kernel holds exactly the plan's gather rows in registers (fewer than
  // the variant's maximum frees registers; the same rows in the same order)
  const int G = std::max(1, std::min(rows, lb_dev_variant_rows(id)));
  auto it = p->spec_kernels.find({id | (G << 8), solver});
  if (it == p->spec_kernels.end()) {
    // no compilation or module load inside a stream capture: a captured
    // first call takes the precompiled kernel (bitwise the same results)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return LB_OK;
    int R, maxw, waves;
    if (!lb_dev_variant(id, &R, &maxw, &waves)) return fail(LB_ERR_INVALID, "no such kernel variant");
    if (!lb_dev_variant_pair(id)) return fail(LB_ERR_INVALID, "only paired variants are specialised");
    char name[96];
    std::snprintf(name, sizeof name, "lb_solve_pair_kernel<%d, %d, %d>", solver == LB_SOLVER_TSIT5 ? 1 : 0, G,
                  waves);
    hipModule_t mod = nullptr;
    hipFunction_t f = nullptr;
    std::string why;
    if (lb_rtc_compile(p->spec_source, name, &mod, &f, why)) {
      // the precompiled kernel computes the same values: run it, remember
      // the failure so that later calls neither recompile nor fail
      (void)hipGetLastError();  // a failed module load leaves its error behind
      mod = nullptr;
      f = nullptr;
      p->spec_failure = std::string(name) + ": " + why;
    }
    it = p->spec_kernels.emplace(std::make_pair(id | (G << 8), solver), std::make_pair(mod, f)).first;
  }
  *fn = it->second.second;
  return LB_OK;
}

}  // namespace

extern "C" int lb_solve(const lb_plan* p, const lb_solve_args* a, void* stream) {
  if (!p || !a) return fail(LB_ERR_INVALID, "null argument");
  if (a->n_traj <= 0) return LB_OK;
  const DevImage* im = nullptr;
  DevCall c{};
  int rc = fill_call(p, a, &im, &c);
  if (rc) return rc;
  int ev_depth = 0;
  if ((rc = prepare_event(const_cast<lb_plan*>(p), a, (hipStream_t)stream, c, &ev_depth))) return rc;
  const bool event = a->event_kind != LB_EVENT_NONE;
  if (event || a->solver >= LB_SOLVER_FIXED_DOPRI5) return launch_table(p, a, im, c, stream, event, ev_depth);
  int id, W;
  const bool pair_ok = a->mode == LB_MODE_PERSISTENT && !a->trace && a->n_saveat == 0 &&
                       !a->replay_h && std::isfinite(a->dt) && a->dt > 0.0 && a->output_kind != LB_OUT_RATE;
  if ((rc = resolve_config(im, a->variant, a->threads_per_traj, pair_ok, &id, &W))) return rc;
  const bool pair = lb_dev_variant_pair(id) != 0;
  void* fn = lb_dev_solve_kernel(id, a->solver);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "kernel variant is not built");
  void* reset = lb_dev_queue_reset_kernel();
  hipStream_t s = (hipStream_t)stream;
  DevPlan P = im->dev;
  const size_t lds = (size_t)(pair ? lds_bytes_pair(P) : im->lds_bytes);
  if ((rc = set_lds_limit(fn, lds))) return rc;
  if (a->mode == LB_MODE_PLAIN) {
    // one launch per trajectory: the per-trajectory GPU path
    int slot = 0;
    if ((rc = const_cast<lb_plan*>(p)->queues.acquire(s, &c.queue, &slot))) return rc;
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
    return const_cast<lb_plan*>(p)->queues.retire(slot, s);
  }
  int grid = a->workgroups;
  if (grid <= 0) {
    int per_cu = 0;
    {
      std::lock_guard<std::mutex> lock(const_cast<lb_plan*>(p)->mu);
      auto& cache = const_cast<DevImage*>(im)->per_cu;
      auto it = cache.find({id, W});
      if (it != cache.end()) {
        per_cu = it->second;
      } else {
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
        cache[{id, W}] = per_cu;
      }
    }
    grid = std::max(1, per_cu) * std::max(1, p->cus);
  }
  grid = (int)std::min<int64_t>(grid, pair ? (a->n_traj + 1) / 2 : a->n_traj);
  hipFunction_t spec_fn = nullptr;
  if (pair && P.time_dependent && lds <= 65536 &&
      (rc = spec_kernel(const_cast<lb_plan*>(p), id, a->solver, (int)im->rows_total, s, &spec_fn)))
    return rc;
  if (spec_fn && a->workgroups <= 0) {  // its own register allocation sets its residency
    int per_cu = 0;
    HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spec_fn, W, lds));
    grid = (int)std::min<int64_t>((int64_t)std::max(1, per_cu) * std::max(1, p->cus),
                                  pair ? (a->n_traj + 1) / 2 : a->n_traj);
  }
  c.traj_offset = 0;
  int slot = 0;
  if ((rc = const_cast<lb_plan*>(p)->queues.acquire(s, &c.queue, &slot))) return rc;
  void* rargs[] = {&c.queue};
  HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
  void* args[] = {&P, &c};
  if (spec_fn)  // dynamic LDS within 64 KB: no per-kernel opt-in needed
    HIP_TRY(hipModuleLaunchKernel(spec_fn, grid, 1, 1, W, 1, 1, (unsigned)lds, s, args, nullptr));
  else
    HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
  HIP_TRY(hipGetLastError());
  return const_cast<lb_plan*>(p)->queues.retire(slot, s);
}

namespace {

// Must match carve_adj() in lb_kernels.hip.
int64_t lds_bytes_adj(const DevPlan& P, int E) {
  const int64_t rows = 1 + 7 * (int64_t)E;
  return 8 * (2 * (int64_t)P.n_pub) + 8 * 2 * (int64_t)P.S + 8 * 2 * (int64_t)P.S +
         8 * (int64_t)((P.n + 1) & ~1) + 16 * rows * (int64_t)(P.KD + 1) + 16 * rows * (int64_t)P.n_slots +
         8 * (int64_t)LB_MAX_DIFF * 16 + 16 * std::max<int64_t>(8, 7 * (int64_t)E) * (int64_t)P.vm_depth + 16;
}

// LDS of the single-block adjoint with `nbuf` segment buffers (carve_aw in
// lb_kernels.hip).
int64_t lds_bytes_adj_wave(const DevPlan& P, const std::vector<int32_t>& g_rows, int t_rows, int every, int nbuf) {
  auto odd = [](int64_t v) { return v | 1; };
  const int64_t g = g_rows.empty() ? 0 : g_rows[0], KC = P.KD + 1;
  return g * 64 * 16 + (int64_t)t_rows * 64 * 16 + KC * 16 + (int64_t)nbuf * every * 7 * odd(KC) * 16 +
         (int64_t)P.n_slots * 16 + 64 * odd(P.n_slots) * 16 + 64 * odd(P.vm_depth) * 16 + 64 * 16 +
         (int64_t)LB_MAX_DIFF * 64 * 16 +
         (int64_t)((P.n + 1) & ~1) * 8 + g * 64 * 4 + (int64_t)t_rows * 64 * 4 + 2 * LB_AW_SWEEPERS * 64 * 16 +
         (4 + (int64_t)LB_MAX_DIFF) * LB_AW_SWEEPERS * 8 + 8 * 4;
}

// The single-block adjoint runs LB_AW_SWEEPERS + 1 waves per trajectory when
// two segment buffers fit in LDS (LB_AW_SWEEPERS sweep a segment while the
// last prepares the next), LB_AW_SWEEPERS otherwise (one buffer; the first
// wave prepares each segment before the sweep).

}  // namespace

// The shared core of lb_adjoint and lb_adjoint_noalloc.  With scratch ==
// nullptr the workspace is the image's grow-only allocation and the queue
// counter comes from the plan's ring; otherwise `scratch` (device,
// scratch_bytes) holds the counter in its first 256 bytes and the workspace
// after it, the grid shrinks to the workgroups whose workspace fits, and
// nothing is allocated or recorded.
int lb_adjoint_launch(const lb_plan* p, const lb_solve_args* a, void* stream, void* scratch,
                      size_t scratch_bytes) {
  if (!p || !a) return fail(LB_ERR_INVALID, "null argument");
  if (a->n_traj <= 0) return LB_OK;
  if (p->src.n_diff <= 0) return fail(LB_ERR_INVALID, "the plan has no differentiated parameters");
  if (a->solver != LB_SOLVER_DOPRI5 && a->solver != LB_SOLVER_TSIT5)
    return fail(LB_ERR_UNSUPPORTED, "the adjoint differentiates the adaptive solvers (dopri5, tsit5)");
  if (!a->tape_h || !a->out_grad) return fail(LB_ERR_INVALID, "lb_adjoint needs the recorded tape and out_grad");
  if (a->output_kind != LB_OUT_INTEGRAL || a->n_weights <= 0)
    return fail(LB_ERR_INVALID, "the adjoint differentiates the weighted integral output");
  if (a->n_saveat > 0) return fail(LB_ERR_UNSUPPORTED, "the adjoint supports pushes at every accepted step only");
  if (a->event_kind != LB_EVENT_NONE) return fail(LB_ERR_UNSUPPORTED, "the adjoint takes no stop event");
  const DevImage* im = nullptr;
  DevCall c{};
  int rc = fill_call(p, a, &im, &c, scratch == nullptr);
  if (rc) return rc;
  c.out_grad = a->out_grad;
  if (a->ckpt_every <= 0) return fail(LB_ERR_INVALID, "ckpt_every must be positive");
  if (a->adjoint_mode != LB_ADJOINT_ADAPTIVE && a->adjoint_mode != LB_ADJOINT_FROZEN)
    return fail(LB_ERR_INVALID, "unknown adjoint mode");
  if (a->adjoint_mode == LB_ADJOINT_ADAPTIVE && !p->src.diff_time)
    return fail(LB_ERR_INVALID,
                "the adjoint of the adaptive solve needs the time derivative programs (lb_plan_desc.diff_time = 1)");
  c.adj_mode = a->adjoint_mode;
  // A plan of one 64-slot block whose gather rows and LDS fit takes the
  // single-block adjoint (lane = slot, several waves per trajectory);
  // otherwise the multi-wave variant with the fewest slots per thread that
  // covers the plan.  args->variant = k + 1 forces adjoint variant k (both
  // forms compute the same derivative).
  const int S = im->dev.S;
  int wave_bufs =
      lds_bytes_adj_wave(im->dev, im->g_rows_host, im->t_rows_host, a->ckpt_every, 2) <= kMaxLdsBytes ? 2 : 1;
  // Polls of one wait between the single-block adjoint's waves: about a
  // second, far beyond the longest legitimate wait (one segment's rows and
  // recomputation, well under a millisecond).  Diagnostic hooks for tests:
  // CENTREX_LB_ADJOINT_BUFFERS=1 forces one segment buffer (no preparing wave),
  // CENTREX_LB_ADJOINT_WAIT_LIMIT sets the bound (negative: every wait counts
  // as run out, so the launch ends with NaN gradients).
  c.adj_wait_limit = 1 << 22;
  if (const char* e = std::getenv("CENTREX_LB_ADJOINT_BUFFERS"))
    if (std::atoi(e) == 1) wave_bufs = 1;
  if (const char* e = std::getenv("CENTREX_LB_ADJOINT_WAIT_LIMIT")) c.adj_wait_limit = std::atoi(e);
  const int wave_threads = 64 * (wave_bufs == 2 ? LB_AW_SWEEPERS + 1 : LB_AW_SWEEPERS);
  const int64_t lds_wave = lds_bytes_adj_wave(im->dev, im->g_rows_host, im->t_rows_host, a->ckpt_every, wave_bufs);
  auto wave_fits = [&](int rows) {
    return im->dev.nblk == 1 && !im->g_rows_host.empty() && im->g_rows_host[0] <= rows &&
           lds_wave <= kMaxLdsBytes;
  };
  int id = -1, W = 0, R = 0, maxw = 0, rows = 0;
  bool wave = false;
  if (a->variant > 0) {
    if (!lb_dev_adjoint_variant(a->variant - 1, &R, &maxw, &rows))
      return fail(LB_ERR_INVALID, "no such adjoint variant");
    wave = rows > 0;
    W = wave ? wave_threads : std::max(64, ((S + R - 1) / R + 63) / 64 * 64);
    if (wave ? !wave_fits(rows) : W > maxw)
      return fail(LB_ERR_INVALID, "the adjoint variant does not cover this plan");
    id = a->variant - 1;
  } else {
    for (int v = 0; lb_dev_adjoint_variant(v, &R, &maxw, &rows); ++v)
      if (rows > 0 && wave_fits(rows)) { id = v; W = wave_threads; wave = true; break; }
    for (int v = 0; id < 0 && lb_dev_adjoint_variant(v, &R, &maxw, &rows); ++v) {
      if (rows > 0) continue;
      const int w = std::max(64, ((S + R - 1) / R + 63) / 64 * 64);
      if (w <= maxw) { id = v; W = w; }
    }
  }
  if (id < 0) return fail(LB_ERR_UNSUPPORTED, "plan is too large for the adjoint kernel variants");
  void* fn = lb_dev_adjoint_kernel(id, a->solver);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "adjoint kernel variant is not built");
  // the multi-wave form evaluates the coefficient rows of E attempts per pass
  // (7E lanes): the largest E of 8, 4, 2 whose LDS keeps the workgroups per CU
  // that E = 1 has (the registers often allow one workgroup per CU anyway)
  int row_batch = 1;
  if (!wave) {
    const int64_t lds1 = lds_bytes_adj(im->dev, 1);
    if ((rc = set_lds_limit(fn, (size_t)lds1))) return rc;
    int base_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&base_cu, fn, W, (size_t)lds1));
    for (int e : {8, 4, 2}) {
      const int64_t le = lds_bytes_adj(im->dev, e);
      if (le > kMaxLdsBytes || 7 * e > W) continue;
      if ((rc = set_lds_limit(fn, (size_t)le))) return rc;
      int per_cu = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, W, (size_t)le));
      if (per_cu >= base_cu && per_cu > 0) {
        row_batch = e;
        break;
      }
    }
  }
  c.adj_row_batch = row_batch;
  const size_t lds = (size_t)(wave ? lds_wave : lds_bytes_adj(im->dev, row_batch));
  if ((rc = set_lds_limit(fn, lds))) return rc;
  int grid = a->workgroups;
  if (grid <= 0) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
    grid = std::max(1, per_cu) * std::max(1, p->cus);
  }
  grid = (int)std::min<int64_t>(grid, a->n_traj);
  // workspace: ckpt_every recomputed attempts per resident workgroup (y and
  // k1..k7, per segment buffer, and the seven coefficient rows of each), in
  // doubles
  const size_t per_y = (size_t)a->ckpt_every * 8 * S * 2 * (wave ? wave_bufs : 1);
  const size_t per_c = (size_t)a->ckpt_every * 7 * (im->dev.KD + 1) * 2;
  hipStream_t s = (hipStream_t)stream;
  if (scratch) {
    const size_t head = 256, per_wg = (per_y + per_c) * sizeof(double);
    if (scratch_bytes < head + per_wg)
      return fail(LB_ERR_UNSUPPORTED, "the no-allocation adjoint needs " + std::to_string(per_wg) +
                                          " bytes of workspace per workgroup (ckpt_every " +
                                          std::to_string(a->ckpt_every) + "); its scratch holds " +
                                          std::to_string(scratch_bytes - head));
    grid = (int)std::min<size_t>((size_t)grid, (scratch_bytes - head) / per_wg);
    char* base = static_cast<char*>(scratch);
    c.queue = reinterpret_cast<int32_t*>(base);
    c.seg_y = reinterpret_cast<double*>(base + head);
    c.seg_c = c.seg_y + (size_t)grid * per_y;
    DevPlan P = im->dev;
    c.traj_offset = 0;
    void* rargs[] = {&c.queue};  // the counter starts at zero (a fill kernel, no memset)
    HIP_TRY(hipLaunchKernel(lb_dev_queue_reset_kernel(), dim3(1), dim3(1), rargs, 0, s));
    void* args[] = {&P, &c};
    HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
    HIP_TRY(hipGetLastError());
    return LB_OK;
  }
  const size_t need_y = (size_t)grid * per_y, need_c = (size_t)grid * per_c;
  bool cap = false;
  if ((rc = is_capturing(s, &cap))) return rc;
  // one adjoint at a time uses an image's workspace: a call waits (on the
  // device) for the previous one, whatever its stream
  std::lock_guard<std::mutex> lock(const_cast<lb_plan*>(p)->mu);
  DevImage* mi = const_cast<DevImage*>(im);
  {
    if (mi->ws_y_len < need_y) {
      if (mi->ws_y) (void)hipFree(mi->ws_y);
      mi->ws_y = nullptr;
      mi->ws_y_len = 0;
      HIP_TRY(hipMalloc(&mi->ws_y, need_y * sizeof(double)));
      mi->ws_y_len = need_y;
    }
    if (mi->ws_c_len < need_c) {
      if (mi->ws_c) (void)hipFree(mi->ws_c);
      mi->ws_c = nullptr;
      mi->ws_c_len = 0;
      HIP_TRY(hipMalloc(&mi->ws_c, need_c * sizeof(double)));
      mi->ws_c_len = need_c;
    }
    c.seg_y = static_cast<double*>(mi->ws_y);
    c.seg_c = static_cast<double*>(mi->ws_c);
  }
  if (!cap && mi->ws_done) HIP_TRY(hipStreamWaitEvent(s, mi->ws_done, 0));
  DevPlan P = im->dev;
  c.traj_offset = 0;
  int slot = 0;
  if ((rc = const_cast<lb_plan*>(p)->queues.acquire(s, &c.queue, &slot))) return rc;
  void* rargs[] = {&c.queue};
  HIP_TRY(hipLaunchKernel(lb_dev_queue_reset_kernel(), dim3(1), dim3(1), rargs, 0, s));
  void* args[] = {&P, &c};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
  HIP_TRY(hipGetLastError());
  if (!cap) {
    if (!mi->ws_done) HIP_TRY(hipEventCreateWithFlags(&mi->ws_done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(mi->ws_done, s));
  }
  return const_cast<lb_plan*>(p)->queues.retire(slot, s);
}

extern "C" int lb_adjoint(const lb_plan* p, const lb_solve_args* a, void* stream) {
  return lb_adjoint_launch(p, a, stream, nullptr, 0);
}

int lb_set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

// Fork streams and events of lb_solve_group, per device, created on first use
// (default priority, non-blocking) and kept until the process ends.
struct GroupStreams {
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;  // [0] fork, [1 + g] join of group g
};
std::mutex g_group_mu;
std::map<int, GroupStreams> g_group;

}  // namespace

extern "C" int lb_solve_group(const lb_plan* const* plans, const lb_solve_args* args, int32_t n_groups,
                              void* stream) {
  if (!plans || !args) return fail(LB_ERR_INVALID, "null argument");
  if (n_groups < 1 || n_groups > LB_MAX_GROUPS) return fail(LB_ERR_INVALID, "n_groups must be 1..16");
  for (int g = 0; g < n_groups; ++g) {
    if (!plans[g]) return fail(LB_ERR_INVALID, "null plan in group " + std::to_string(g));
    if (args[g].mode != LB_MODE_PERSISTENT)
      return fail(LB_ERR_INVALID, "grouped solves use the persistent mode");
    if (args[g].event_kind != LB_EVENT_NONE)
      return fail(LB_ERR_UNSUPPORTED, "grouped solves take no stop events (solve each plan with lb_solve)");
    if (plans[g]->device != plans[0]->device) return fail(LB_ERR_INVALID, "all plans must be on one device");
  }
  if (n_groups == 1) return lb_solve(plans[0], &args[0], stream);
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_group_mu);
  GroupStreams& gs = g_group[plans[0]->device];
  while ((int)gs.streams.size() < n_groups) {
    hipStream_t t;
    HIP_TRY(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    gs.streams.push_back(t);
  }
  while ((int)gs.events.size() < n_groups + 1) {
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    gs.events.push_back(e);
  }
  // fork: every group stream starts after the work already queued on `stream`
  HIP_TRY(hipEventRecord(gs.events[0], s));
  for (int g = 0; g < n_groups; ++g) HIP_TRY(hipStreamWaitEvent(gs.streams[g], gs.events[0], 0));
  int rc = LB_OK;
  for (int g = 0; g < n_groups && rc == LB_OK; ++g) rc = lb_solve(plans[g], &args[g], gs.streams[g]);
  // join: `stream` continues after every group (also after a failed launch,
  // so a capture stays balanced)
  for (int g = 0; g < n_groups; ++g) {
    HIP_TRY(hipEventRecord(gs.events[1 + g], gs.streams[g]));
    HIP_TRY(hipStreamWaitEvent(s, gs.events[1 + g], 0));
  }
  return rc;
}

extern "C" int lb_launch_info(const lb_plan* p, const uint8_t* y0_support, int32_t variant,
                              int32_t threads_per_traj, int32_t* out) {
  if (!p || !out) return fail(LB_ERR_INVALID, "null argument");
  const DevImage* im = nullptr;
  int rc = select_image(const_cast<lb_plan*>(p), y0_support, &im);
  if (rc) return rc;
  int id, W;
  if ((rc = resolve_config(im, variant, threads_per_traj, true, &id, &W))) return rc;
  void* fn = lb_dev_solve_kernel(id, LB_SOLVER_DOPRI5);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "kernel variant is not built");
  const size_t lds = (size_t)(lb_dev_variant_pair(id) ? lds_bytes_pair(im->dev) : im->lds_bytes);
  if ((rc = set_lds_limit(fn, lds))) return rc;
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
  out[0] = id + 1;
  out[1] = W;
  out[2] = per_cu;
  out[3] = p->cus;
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

// ===========================================================================
// Effective-Hamiltonian plans (rust/src/effective_lindblad/plan.rs:122-190,
// solver.rs:78-281).
// ===========================================================================

extern "C" int lb_dev_eff_variant(int id, int* R, int* maxw);
extern "C" void* lb_dev_eff_kernel(int id, int solver);
extern "C" void* lb_dev_eff_rhs_kernel(int id);

struct lb_eff_plan {
  DevEffPlan dev{};
  std::vector<void*> allocations;
  int device = 0, cus = 0;
  QueueRing queues;
  ~lb_eff_plan() {
    for (void* a : allocations) (void)hipFree(a);
  }
};

namespace {

// Must match carve_eff() in lb_kernels.hip.
int64_t lds_bytes_eff(const DevEffPlan& P) {
  return 8 * 2 * (int64_t)P.S + 16 * 8 * (int64_t)P.n_slots + 16 * 8 * (int64_t)P.vm_depth +
         8 * 4 * (int64_t)P.nblk + 8 * 32 + 8 * 2 + 4 * 4;
}

template <typename T>
int eff_upload(lb_eff_plan* p, const std::vector<T>& host, const T** out) {
  void* d = nullptr;
  HIP_TRY(hipMalloc(&d, std::max<size_t>(host.size(), 1) * sizeof(T)));
  p->allocations.push_back(d);
  if (!host.empty()) HIP_TRY(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(d);
  return LB_OK;
}

// The multi-wave variant with the fewest rows per lane that covers D rows.
int eff_config(const DevEffPlan& P, int* id, int* W) {
  int R = 0, maxw = 0;
  for (int v = 0; lb_dev_eff_variant(v, &R, &maxw); ++v) {
    const int w = std::max(64, ((P.D + R - 1) / R + 63) / 64 * 64);
    if (w <= maxw && (int64_t)R * w >= P.D) {
      *id = v;
      *W = w;
      return LB_OK;
    }
  }
  return fail(LB_ERR_UNSUPPORTED, "effective plan has more rows than the kernel variants cover (4096)");
}

}  // namespace

extern "C" int lb_eff_plan_create(const lb_eff_plan_desc* d, lb_eff_plan** out_plan) {
  if (!out_plan) return fail(LB_ERR_INVALID, "out_plan is null");
  *out_plan = nullptr;
  if (!d || d->n_states <= 0) return fail(LB_ERR_INVALID, "n_states must be positive");
  const int n = d->n_states, D = d->real_dim;
  if ((int64_t)D != 2 * (int64_t)n * n) return fail(LB_ERR_INVALID, "real_dim must be 2 * n_states^2");
  if (d->n_grid < 2 || !d->field_grid) return fail(LB_ERR_INVALID, "the operator grid needs at least two points");
  if (d->n_base < 0 || d->n_base > d->n_slots) return fail(LB_ERR_INVALID, "n_base exceeds n_slots");
  for (int s : {d->field_slot, d->rabi_slot, d->detuning_slot})
    if (s < 0 || s >= d->n_slots) return fail(LB_ERR_INVALID, "field, Rabi or detuning slot out of range");
  std::unique_ptr<lb_eff_plan> p(new lb_eff_plan());
  DevEffPlan& P = p->dev;
  P.n = n;
  P.D = D;
  P.G = d->n_grid;
  P.S = (D + 63) / 64 * 64;
  P.nblk = P.S / 64;
  P.field_slot = d->field_slot;
  P.rabi_slot = d->rabi_slot;
  P.det_slot = d->detuning_slot;
  P.n_slots = d->n_slots;
  P.n_base = d->n_base;
  int rc;
  if ((rc = eff_upload(p.get(), std::vector<double>(d->field_grid, d->field_grid + d->n_grid), &P.grid))) return rc;
  for (int o = 0; o < 3; ++o) {
    const int64_t nnz = d->nnz[o];
    if (nnz < 0 || nnz > INT32_MAX / 4) return fail(LB_ERR_UNSUPPORTED, "operator has too many stored entries");
    if (!d->row_ptrs[o] || (nnz > 0 && (!d->col_indices[o] || !d->coeffs[o])))
      return fail(LB_ERR_INVALID, "operator tables are required");
    std::vector<int32_t> rp(D + 1), col(nnz);
    if (d->row_ptrs[o][0] != 0 || d->row_ptrs[o][D] != nnz)
      return fail(LB_ERR_INVALID, "operator row_ptrs must start at 0 and end at nnz");
    for (int i = 0; i <= D; ++i) {
      if (i < D && d->row_ptrs[o][i] > d->row_ptrs[o][i + 1])
        return fail(LB_ERR_INVALID, "operator row_ptrs must be monotonic");
      rp[i] = (int32_t)d->row_ptrs[o][i];
    }
    for (int64_t q = 0; q < nnz; ++q) {
      if (d->col_indices[o][q] < 0 || d->col_indices[o][q] >= D)
        return fail(LB_ERR_INVALID, "operator column index out of range");
      col[q] = (int32_t)d->col_indices[o][q];
    }
    // sliced layout (lb_device.h): per 64-row slice, its longest row's
    // entries x 64 cells; padding cells are never read
    const int S = P.S, nblk = P.nblk, GI = d->n_grid - 1;
    std::vector<int32_t> len(S, 0);
    std::vector<int64_t> sp(nblk);
    int64_t cells = 0;
    for (int b = 0; b < nblk; ++b) {
      int longest = 0;
      for (int r = 64 * b; r < std::min(D, 64 * b + 64); ++r) {
        len[r] = rp[r + 1] - rp[r];
        longest = std::max(longest, len[r]);
      }
      sp[b] = cells;
      cells += 64 * (int64_t)longest;
    }
    if (cells * (int64_t)std::max(GI, 1) > ((int64_t)1 << 31))
      return fail(LB_ERR_UNSUPPORTED, "operator tables too large in sliced form");
    std::vector<int32_t> scol(std::max<int64_t>(cells, 1), 0);
    std::vector<double2> lo((size_t)std::max<int64_t>(cells * GI, 1), double2{0.0, 0.0}), hi(lo);
    for (int r = 0; r < D; ++r) {
      for (int j = 0; j < len[r]; ++j) {
        const int64_t q = rp[r] + j, c = sp[r / 64] + 64 * (int64_t)j + r % 64;
        scol[c] = col[q];
        for (int k = 0; k < GI; ++k) {
          const double* e = d->coeffs[o] + ((size_t)k * nnz + q) * 4;
          lo[(size_t)k * cells + c] = double2{e[0], e[1]};
          hi[(size_t)k * cells + c] = double2{e[2], e[3]};
        }
      }
    }
    P.cells[o] = cells;
    if ((rc = eff_upload(p.get(), len, &P.row_len[o]))) return rc;
    if ((rc = eff_upload(p.get(), sp, &P.slice_ptr[o]))) return rc;
    if ((rc = eff_upload(p.get(), scol, &P.col[o]))) return rc;
    if ((rc = eff_upload(p.get(), lo, &P.coef_lo[o]))) return rc;
    if ((rc = eff_upload(p.get(), hi, &P.coef_hi[o]))) return rc;
  }
  // parameter graph: base values and compounds (checked as plan expressions)
  std::vector<double> base(2 * (size_t)std::max(d->n_base, 0));
  for (int i = 0; i < d->n_base; ++i) {
    base[2 * i] = d->base_re[i];
    base[2 * i + 1] = d->base_im[i];
  }
  if ((rc = eff_upload(p.get(), base, &P.base))) return rc;
  lb_plan_desc g{};
  g.instr_op = d->instr_op;
  g.instr_index = d->instr_index;
  g.instr_argc = d->instr_argc;
  g.instr_function = d->instr_function;
  g.instr_re = d->instr_re;
  g.instr_im = d->instr_im;
  std::vector<DevOp> ops;
  std::vector<DevExpr> comp;
  int depth = 1;
  for (int c = 0; c < d->n_compounds; ++c) {
    const int a = d->expr_ptrs[c], b = d->expr_ptrs[c + 1];
    if (d->compound_slots[c] < d->n_base || d->compound_slots[c] >= d->n_slots)
      return fail(LB_ERR_INVALID, "compound slot out of range");
    std::string why;
    int dep = 0;
    const int r = check_expression(&g, a, b, d->n_slots, why, &dep);
    if (r == -1) return fail(LB_ERR_INVALID, "compound " + std::to_string(c) + ": " + why);
    if (r == -2) return fail(LB_ERR_UNSUPPORTED, "compound " + std::to_string(c) + ": " + why);
    depth = std::max(depth, dep);
    DevExpr e{};
    e.start = (int32_t)ops.size();
    for (int q = a; q < b; ++q) {
      DevOp op{};
      op.code = (d->instr_op[q] & 0xFF) | ((d->instr_argc[q] & 0xFF) << 8) | ((d->instr_function[q] & 0xFFFF) << 16);
      op.index = d->instr_index[q];
      op.re = d->instr_re[q];
      op.im = d->instr_im[q];
      op.aux = 0.0;
      ops.push_back(op);
    }
    e.end = (int32_t)ops.size();
    e.dst = d->compound_slots[c];
    comp.push_back(e);
  }
  P.n_comp = (int32_t)comp.size();
  P.vm_depth = depth;
  if ((rc = eff_upload(p.get(), ops, &P.instr))) return rc;
  if ((rc = eff_upload(p.get(), comp, &P.comp))) return rc;
  HIP_TRY(hipGetDevice(&p->device));
  HIP_TRY(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device));
  if ((rc = p->queues.init())) return rc;
  int id, W;
  if ((rc = eff_config(P, &id, &W))) return rc;
  if (lds_bytes_eff(P) > kMaxLdsBytes) return fail(LB_ERR_UNSUPPORTED, "effective plan needs more LDS than a workgroup has");
  *out_plan = p.release();
  return LB_OK;
}

extern "C" void lb_eff_plan_destroy(lb_eff_plan* p) { delete p; }

extern "C" int lb_eff_solve(const lb_eff_plan* p, const lb_eff_solve_args* a, void* stream) {
  if (!p || !a) return fail(LB_ERR_INVALID, "null argument");
  if (a->n_traj <= 0) return LB_OK;
  if (a->solver != LB_SOLVER_DOPRI5 && a->solver != LB_SOLVER_TSIT5)
    return fail(LB_ERR_UNSUPPORTED, "the effective-Hamiltonian GPU path runs dopri5 and tsit5");
  if (!a->y0 || !a->out_values || !a->out_stats || !a->out_fail_t || !a->out_counts)
    return fail(LB_ERR_INVALID, "y0 and output buffers are required");
  if (a->maxiters <= 0) return fail(LB_ERR_INVALID, "maxiters must be positive");
  if (!std::isfinite(a->t0) || !std::isfinite(a->t1)) return fail(LB_ERR_INVALID, "t_span must be finite");
  if (a->capacity <= 0) return fail(LB_ERR_INVALID, "capacity must be positive");
  const DevEffPlan& P = p->dev;
  DevEffCall c{};
  c.n_traj = a->n_traj;
  c.t0 = a->t0;
  c.t1 = a->t1;
  c.abstol = a->abstol;
  c.reltol = a->reltol;
  c.dt = a->dt;
  c.maxiters = a->maxiters;
  c.save_start = a->save_start;
  c.y0 = a->y0;
  if (a->n_override < 0 || a->n_override > LB_MAX_OVERRIDES)
    return fail(LB_ERR_UNSUPPORTED, "at most 16 overridden parameters");
  if (a->n_override > 0 && !a->override_values) return fail(LB_ERR_INVALID, "override_values are required");
  for (int j = 0; j < a->n_override; ++j) {
    if (a->override_slots[j] < 0 || a->override_slots[j] >= P.n_slots)
      return fail(LB_ERR_INVALID, "parameter override slot " + std::to_string(a->override_slots[j]) + " out of range");
    c.override_slot[j] = a->override_slots[j];
  }
  c.n_override = a->n_override;
  c.override_values = a->override_values;
  c.output_kind = a->output_kind;
  switch (a->output_kind) {
    case LB_EFF_OUT_POPULATIONS: c.width = P.n; break;
    case LB_EFF_OUT_FULL: c.width = P.D; break;
    case LB_EFF_OUT_INTEGRAL: c.width = 1; break;
    case LB_EFF_OUT_SELECTED: {
      if (a->n_selected <= 0 || a->n_selected > LB_MAX_SELECTED)
        return fail(LB_ERR_INVALID, "output='selected' requires 1..128 output_indices");
      const int n = P.n, n2 = n * n;
      for (int q = 0; q < a->n_selected; ++q) {
        const int i = a->selected_rows[q], j = a->selected_cols[q];
        if (i < 0 || j < 0 || i >= n || j >= n) return fail(LB_ERR_INVALID, "selected index out of range");
        // split_real_selected_extractions (solver.rs:60-76)
        c.sel_re[q] = i * n + j;
        c.sel_im[q] = i == j ? -1 : n2 + i * n + j;
      }
      c.n_selected = a->n_selected;
      c.width = 2 * a->n_selected;
      break;
    }
    default: return fail(LB_ERR_INVALID, "unknown output kind");
  }
  if (a->output_kind == LB_EFF_OUT_INTEGRAL) {
    if (a->n_weights <= 0 || a->n_weights > LB_MAX_WEIGHTS)
      return fail(LB_ERR_INVALID, "integral outputs require 1..64 integral weights");
    for (int q = 0; q < a->n_weights; ++q) {
      if (a->weight_indices[q] < 0 || a->weight_indices[q] >= P.D)
        return fail(LB_ERR_INVALID, "integral weight index out of range");
      c.weight_index[q] = a->weight_indices[q];
      c.weight_value[q] = a->weight_values[q];
    }
    c.n_weights = a->n_weights;
  }
  if (a->n_saveat < 0 || (a->n_saveat > 0 && !a->saveat)) return fail(LB_ERR_INVALID, "bad save plan");
  c.n_saveat = a->n_saveat;
  c.saveat = a->saveat;
  c.capacity = a->capacity;
  c.out_values = a->out_values;
  c.out_times = a->out_times;
  c.out_counts = a->out_counts;
  c.out_stats = a->out_stats;
  c.out_fail_t = a->out_fail_t;
  int id, W, rc;
  if ((rc = eff_config(P, &id, &W))) return rc;
  void* fn = lb_dev_eff_kernel(id, a->solver);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "effective kernel variant is not built");
  const size_t lds = (size_t)lds_bytes_eff(P);
  if ((rc = set_lds_limit(fn, lds))) return rc;
  hipStream_t s = (hipStream_t)stream;
  int grid = a->workgroups;
  if (grid <= 0) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, W, lds));
    grid = std::max(1, per_cu) * std::max(1, p->cus);
  }
  grid = (int)std::min<int64_t>(grid, a->n_traj);
  void* reset = lb_dev_queue_reset_kernel();
  int slot = 0;
  if ((rc = const_cast<lb_eff_plan*>(p)->queues.acquire(s, &c.queue, &slot))) return rc;
  void* rargs[] = {&c.queue};
  HIP_TRY(hipLaunchKernel(reset, dim3(1), dim3(1), rargs, 0, s));
  DevEffPlan Pc = P;
  void* args[] = {&Pc, &c};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, s));
  HIP_TRY(hipGetLastError());
  return const_cast<lb_eff_plan*>(p)->queues.retire(slot, s);
}

extern "C" int lb_eff_rhs_eval(const lb_eff_plan* p, int64_t batch, const double* y, const double* t,
                               double* out, int32_t* out_status, void* stream) {
  if (!p || !y || !t || !out || !out_status) return fail(LB_ERR_INVALID, "null argument");
  if (batch <= 0) return LB_OK;
  const DevEffPlan& P = p->dev;
  int id, W, rc;
  if ((rc = eff_config(P, &id, &W))) return rc;
  void* fn = lb_dev_eff_rhs_kernel(id);
  if (!fn) return fail(LB_ERR_UNSUPPORTED, "effective kernel variant is not built");
  const size_t lds = (size_t)lds_bytes_eff(P);
  if ((rc = set_lds_limit(fn, lds))) return rc;
  DevEffCall c{};
  DevEffRhsCall Q{batch, y, t, out, out_status};
  DevEffPlan Pc = P;
  void* args[] = {&Pc, &c, &Q};
  const int grid = (int)std::min<int64_t>(batch, 4096);
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(W), args, lds, (hipStream_t)stream));
  HIP_TRY(hipGetLastError());
  return LB_OK;
}
