// Device-side data layout of a Lindblad plan and of one solve call.
//
// Lane ownership.  The n(n+1)/2 upper-triangle outputs of the RHS program
// (centrex_tlf/lindblad/ir.py:447-613) are permuted into S "slots", S a
// multiple of 64.  A workgroup of W threads (W a multiple of 64) owns one
// trajectory; thread t holds slots t, t+W, t+2W, ... in registers, so slot
// s = lane + 64*(wave + (W/64)*r).  Each group of 64 consecutive slots (a
// "block") is owned by one wave for one r; reductions sum a block with a
// fixed butterfly and then the blocks in index order, so every result is
// independent of W.
//
// Each slot carries one complex element of rho (the diagonal slots only use
// the real part).  Per RHS evaluation an output is
//   k_s = a_s * x_s + sum_g f_g * c_g(t) * x_{in(g)}
// where a_s folds the time-independent terms that read the slot's own element
// (computed once per trajectory), and the "gather" terms read other elements
// that the owning threads publish to LDS.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else  // runtime compilation (lb_rtc.cpp): fixed-width types of the compiler's prelude
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint64_t uint64_t;
#endif

#include "../../include/centrex_lindblad.h"

#define LB_MAX_OVERRIDES 16
#define LB_MAX_AXES 8
#define LB_MAX_WEIGHTS 64
#define LB_MAX_SELECTED 128
#define LB_MAX_STACK 1024
#define LB_MAX_DIFF 8
#define LB_STATIC_COEF 255u

// gather word: bits 0..19 LDS publication index, bit 20 real input (diagonal),
// bit 21 conjugated input, bit 22 conjugated coefficient, bits 24..31 coefficient id
// (LB_STATIC_COEF = constant 1).
__host__ __device__ inline uint32_t lb_gword(uint32_t pub, bool real_in, bool conj_in, bool conj_c,
                                             uint32_t coef) {
  return (pub & 0xFFFFFu) | (real_in ? (1u << 20) : 0u) | (conj_in ? (1u << 21) : 0u) |
         (conj_c ? (1u << 22) : 0u) | (coef << 24);
}

// One bytecode instruction (rust/src/lindblad/eval.rs:11-60 opcodes).  Op 20
// (POWC, device only) raises the top of the stack to the real constant `re`,
// with im = cos(re*pi) and aux = sin(re*pi) for negative real bases.
#define LB_OP_POWC 20
struct DevInstr {
  int32_t op, index, argc, function;
  double re, im, aux;
};

// The device record of one instruction: 32 bytes, read by the expression
// machine with one scalar load (code = op | argc << 8 | function << 16).
struct alignas(32) DevOp {
  int32_t code;
  int32_t index;
  double re, im, aux;
};
static_assert(sizeof(DevOp) == 32, "one 32-byte record per instruction");

// One expression of the program: instructions [start, end) and where the
// value goes: dst >= 0 a compound's slot, dst < 0 coefficient row entry -dst-1.
struct alignas(16) DevExpr {
  int32_t start, end, dst, pad;
};

struct DevPlan {
  int32_t n, dim, upper_len;
  int32_t S;        // padded slot count (multiple of 64)
  int32_t nblk;     // S / 64
  int32_t n_pub;    // published elements
  int32_t K;        // Hamiltonian coefficients
  int32_t KD;       // coefficient programs: K, plus K * n_diff derivative programs
  int32_t n_diff;   // derivative columns per coefficient (0 for an ordinary plan)
  int32_t diff_time; // 1: the last derivative column is d/dt (the others are parameters)
  int32_t n_slots, n_base, n_compounds;
  int32_t max_self; // self terms per slot (ELL width)
  int32_t time_dependent;
  // per slot
  const int32_t* slot_up;    // upper index or -1 (padding)
  const int32_t* slot_re;    // packed real index or -1
  const int32_t* slot_im;    // packed imag index or -1 (diagonal / padding)
  const int32_t* slot_pub;   // publication index or -1
  const int32_t* self_coef;  // S*max_self, coefficient id or -1 (constant) ; -2 unused
  const uint8_t* self_conj;  // S*max_self
  const double* self_f;      // S*max_self*2
  // gather ELL, rows grouped per block
  const int32_t* blk_row0;   // nblk
  const int32_t* blk_rows;   // nblk
  const uint32_t* g_word;    // rows*64
  const double* g_f;         // rows*64*2
  int32_t g_rows;            // total gather rows
  int32_t lds_tables;        // 1: kernels copy the gather tables into LDS
  // restricted plans hold only the elements reachable from an initial support;
  // the others are written as zeros
  int32_t restricted;
  const uint8_t* upper_active;   // upper_len
  const int32_t* packed_upper;   // dim: packed index -> upper index
  // differentiable plans: for each input slot, the terms reading it (ELL per
  // 64-slot block; word as lb_gword with the OUTPUT slot in bits 0..19)
  int32_t t_rows_total;
  const int32_t* t_blk_row0;
  const int32_t* t_blk_rows;
  const uint32_t* t_word;
  const double* t_f;
  // expression machine
  const DevOp* instr;
  const int32_t* expr_start;    // n_compounds + K + 1
  const int32_t* compound_slot; // n_compounds
  const int32_t* expr_dyn;      // n_compounds + K: 1 if time-dependent
  // expression records in evaluation order: every expression (the solver's
  // n_compounds + K, then the derivative programs), and the time-dependent
  // ones only, with the solver's count first (prefixes of each list)
  const DevExpr* ex_all;
  const DevExpr* ex_dyn;
  int32_t n_all_fwd, n_all, n_dyn_fwd, n_dyn;
  int32_t vm_depth;             // deepest expression stack
  const double* base;           // n_base*2
};

// Waves that sweep a checkpoint segment of the single-block adjoint (each
// applies its share of the transposed rows of every reverse stage).
#define LB_AW_SWEEPERS 4

struct DevCall {
  int64_t n_traj;
  double t0, t1, abstol, reltol, dt;
  int64_t maxiters;
  int32_t save_start;
  int32_t solver;
  const double* t1_per_traj;
  const double* y0;
  int32_t y0_shared;
  int32_t n_override;
  int32_t override_slot[LB_MAX_OVERRIDES];
  const double* override_values;  // n_traj x n_override x 2, or null for a grid
  int32_t n_axes;
  int64_t axis_len[LB_MAX_AXES];
  int64_t axis_stride[LB_MAX_AXES];
  int64_t axis_offset[LB_MAX_AXES];
  const double* axis_values;
  int32_t output_kind;
  int32_t n_weights;
  int32_t weight_index[LB_MAX_WEIGHTS];
  double weight_value[LB_MAX_WEIGHTS];
  int32_t n_selected;
  int32_t sel_upper[LB_MAX_SELECTED];  // upper index of the stored element
  int32_t sel_flags[LB_MAX_SELECTED];  // bit0 diagonal, bit1 conjugate
  double* out_values;
  int64_t* out_stats;
  double* out_fail_t;
  double* out_integral;    // optional trapezoid next to a non-integral output
  int32_t want_integral;   // accumulate the trapezoid
  int32_t* queue;          // work-queue counter
  int64_t traj_offset;     // plain mode: first trajectory of this launch
  int64_t n_total;         // trajectories of the call (all launches of plain mode)
  const int32_t* queue_order;  // queue position -> trajectory, or NULL
  // pushed points: every accepted step (n_saveat == 0) or a save plan
  int32_t n_saveat;
  const double* saveat;    // device, n_saveat ascending times (t0 first when saving the start)
  int32_t trace;           // 1: one output row per pushed point
  int64_t capacity;        // rows per trajectory in trace mode
  double* out_times;       // device n_traj x capacity, or null
  int32_t* out_counts;     // device n_traj (trace mode)
  // tape for the adjoint, one record per step attempt: start time, step size,
  // (error norm, proposed next step, fac_old), LB_TAPE_* flags, and the state
  // at the start of every ckpt_every-th attempt (slot order, re/im)
  double* tape_t;          // n_traj x tape_cap
  double* tape_h;          // n_traj x tape_cap
  double* tape_c;          // n_traj x tape_cap x 3
  int32_t* tape_f;         // n_traj x tape_cap
  double* tape_y;          // n_traj x ceil(tape_cap / ckpt_every) x S x 2
  int32_t* tape_n;         // n_traj attempts recorded
  int64_t tape_cap;
  int32_t ckpt_every;
  int32_t adj_mode;        // LB_ADJOINT_*
  int32_t adj_wait_limit;  // polls of one exchange wait between the waves of the single-block
                           // adjoint before the launch is marked broken (< 0: broken from the start)
  int32_t adj_row_batch;   // multi-wave adjoint: attempts whose coefficient rows one pass evaluates
  // adjoint workspace (per workgroup) and output
  double* seg_y;           // workgroups x ckpt_every x 8 x S x 2: y and k1..k7 of each attempt
  double* seg_c;           // workgroups x ckpt_every x 7 x (KD + 1) x 2: coefficient rows
  double* out_grad;        // n_traj x (n_diff - diff_time): d(integral)/d(parameter)
  // replay: take the accepted step sizes of a recorded tape (stride tape_cap)
  // instead of the step-size controller; every step is accepted
  const double* replay_h;
  const int32_t* replay_n;
  // terminal stop event: population sum (event_idx) or expression (event_ops)
  int32_t event_kind;       // LB_EVENT_*
  int32_t n_event;          // indices or instructions
  const DevOp* event_ops;   // device
  const int32_t* event_idx; // device, diagonal indices in summation order
  double event_threshold;
  double* out_event_t;      // device n_traj: root time, NaN when none fired; or null
};

// ---- effective-Hamiltonian plans (split-real state, interpolated operators)
struct DevEffPlan {
  int32_t n, D, G;          // states, real dimension, grid points
  int32_t S;                // D padded to a multiple of 64
  int32_t nblk;             // S / 64
  const double* grid;       // G
  // Each operator in sliced form: the 64 rows of a wave are one slice, and
  // entry j of a slice's rows lies at cell slice_ptr + 64 j + (row mod 64),
  // so a wave's lanes read consecutive cells; rows keep their CSR order.
  const int32_t* row_len[3];    // S (0 on padding rows)
  const int64_t* slice_ptr[3];  // nblk
  const int32_t* col[3];        // cells
  const double2* coef_lo[3];    // (G - 1) x cells: (c0, c1) of the cubic
  const double2* coef_hi[3];    // (G - 1) x cells: (c2, c3)
  int64_t cells[3];
  int32_t field_slot, rabi_slot, det_slot;
  int32_t n_slots, n_base, n_comp, vm_depth;
  const double* base;       // n_base x 2
  const DevOp* instr;
  const DevExpr* comp;      // compound records in evaluation order (dst = slot)
};

struct DevEffCall {
  int64_t n_traj;
  double t0, t1, abstol, reltol, dt;
  int64_t maxiters;
  int32_t save_start;
  const double* y0;
  int32_t n_override;
  int32_t override_slot[LB_MAX_OVERRIDES];
  const double* override_values;  // n_traj x n_override (real)
  int32_t output_kind;            // LB_EFF_OUT_*
  int32_t width;
  int32_t n_selected;
  int32_t sel_re[LB_MAX_SELECTED];  // index of the real part in y
  int32_t sel_im[LB_MAX_SELECTED];  // index of the imaginary part, or -1 (diagonal)
  int32_t n_weights;
  int32_t weight_index[LB_MAX_WEIGHTS];
  double weight_value[LB_MAX_WEIGHTS];
  int32_t n_saveat;
  const double* saveat;
  int64_t capacity;
  double* out_values;
  double* out_times;
  int32_t* out_counts;
  int64_t* out_stats;
  double* out_fail_t;
  int32_t* queue;
};

struct DevEffRhsCall {
  int64_t batch;
  const double* y;
  const double* t;
  double* out;
  int32_t* status;
};

struct DevRhsCall {
  int64_t batch;
  const double* y;
  const double* t;
  int32_t n_override;
  int32_t override_slot[LB_MAX_OVERRIDES];
  const double* override_values;
  double* out;
};
