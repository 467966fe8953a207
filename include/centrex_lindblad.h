/*
 * C ABI of the MI355X Lindblad / optical-Bloch solver.
 *
 * This header replaces the PyO3 module `centrex_tlf.centrex_tlf_rust`
 * (registered at rust/src/lindblad/mod.rs:18-30).  Every entry point takes
 * plain pointers and sizes, returns an int status (LB_OK on success) and
 * leaves a message for lb_last_error() on failure; no C++ exception crosses
 * it.  Array arguments marked "device" are HIP device pointers (for example
 * torch tensors on a ROCm device); "host" arrays are ordinary memory.
 *
 * Packed state layout (n states, dim = n*n doubles): the n real diagonal
 * entries, then (re, im) of the strict upper triangle in row-major order
 * (centrex_tlf/lindblad/state_layout.py:11-79).
 */
#ifndef CENTREX_LINDBLAD_H
#define CENTREX_LINDBLAD_H

#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define LB_OK 0
#define LB_ERR_INVALID 1     /* bad argument or plan description            */
#define LB_ERR_DEVICE 2      /* HIP runtime failure                          */
#define LB_ERR_UNSUPPORTED 3 /* valid request this build does not implement */

/* Per-trajectory status written to lb_solve_args.out_status. */
#define LB_TRAJ_OK 0
#define LB_TRAJ_MAXITERS 1  /* "Stopped at x=... Need more than N steps."  rust/src/ode/dopri5.rs:226-228 */
#define LB_TRAJ_UNDERFLOW 2 /* "Stopped at x=... Step size underflow."     dopri5.rs:229-231 */
#define LB_TRAJ_STIFF 3     /* "Stiff at x=..."                             dopri5.rs:255-277 */
#define LB_TRAJ_BAD_SPAN 4  /* t1 < t0 ("only forward integration")         dopri5.rs:175-177 */
#define LB_TRAJ_BAD_SUPPORT 5 /* y0 is nonzero outside lb_solve_args.y0_support (no reference
                                 counterpart: the mask is a GPU-side promise) */
#define LB_TRAJ_SAVEAT 6      /* "ended before all saveat: ..."            dopri5.rs:369-376;
                                 out_fail_t holds the number of plan points reached */
#define LB_TRAJ_CAPACITY 7    /* more trace rows than lb_solve_args.capacity (the integration
                                 completed; counts and stats are exact, rows past capacity
                                 were dropped).  No reference counterpart. */
#define LB_TRAJ_BAD_DT 8      /* fixed-step solver without a positive finite dt
                                 ("fixed-step solvers require positive finite dt", fixed.rs:154-156) */
#define LB_TRAJ_EVENT_COMPLEX 9 /* the stop-event expression evaluated to a complex value
                                   ("event expression evaluated to a complex value",
                                   rust/src/lindblad/rhs.rs:1230-1236); fail_t = its time */

/* Terminal stop events (rust/src/lindblad/ode_impl.rs:8-16, 134-153). */
#define LB_EVENT_NONE 0
#define LB_EVENT_POPULATION 1 /* g = sum of rho_kk over event_indices - event_threshold */
#define LB_EVENT_EXPRESSION 2 /* g = Re of the event bytecode over the parameter slots at t */

/* Solvers (rust/src/ode/batch.rs:7-27). */
#define LB_SOLVER_DOPRI5 0
#define LB_SOLVER_TSIT5 1
/* fixed-step solvers of rust/src/ode/fixed.rs:133-249: steps of dt, shortened
 * to land on t1 and on save-plan points, no error control */
#define LB_SOLVER_FIXED_DOPRI5 2
#define LB_SOLVER_FIXED_RK2 3
#define LB_SOLVER_FIXED_RK4 4

/* Output kinds (rust/src/ode/output.rs). */
#define LB_OUT_POPULATIONS 0 /* final diagonal, width n                          */
#define LB_OUT_SELECTED 1    /* final selected entries, complex, width = count   */
#define LB_OUT_INTEGRAL 2    /* trapezoid of sum_k w_k rho_kk over pushed points */
#define LB_OUT_FULL 3        /* final packed state, width dim                     */
#define LB_OUT_RATE 4        /* sum_k w_k rho_kk at each pushed point (trace only;
                                output.rs:337-384 WeightedRateOutput), width 1    */

/* Launch modes. */
#define LB_MODE_PERSISTENT 0  /* one launch, persistent workgroups + work queue */
#define LB_MODE_PLAIN 1       /* one launch per trajectory                      */

/*
 * Plan description: the lowered OBE system.  Mirrors the payload of
 * prepare_lindblad_problem_py (rust/src/lindblad/python_api.rs:271-274,
 * parsed by rust/src/lindblad/plan.rs:1155-1308):
 *   - parameter graph: base slot values, compounds in evaluation order;
 *   - decomposed Hamiltonian coefficients as bytecode (ops 1-19, builtins
 *     1-7, helpers 1-16 of rust/src/lindblad/eval.rs:11-153);
 *   - the expanded sparse RHS program over the n(n+1)/2 upper outputs
 *     (centrex_tlf/lindblad/ir.py:447-613, rust/src/lindblad/plan.rs:649-662).
 * Expressions are concatenated: first the n_compounds compound expressions,
 * then the n_coefficients Hamiltonian coefficients; expr_ptrs has
 * n_compounds + n_coefficients + 1 entries into the instr_* arrays.
 * All arrays are host memory and are copied by lb_plan_create.
 */
typedef struct lb_plan_desc {
  int32_t n_states;
  int32_t n_slots;
  int32_t n_base;
  const double* base_re; /* n_base */
  const double* base_im;
  int32_t n_compounds;
  const int32_t* compound_slots; /* n_compounds */
  int32_t n_coefficients;
  const int32_t* expr_ptrs;
  const int32_t* instr_op;
  const int32_t* instr_index;
  const int32_t* instr_argc;
  const int32_t* instr_function;
  const double* instr_re;
  const double* instr_im;
  int64_t n_terms;
  const int64_t* output_ptrs;   /* upper_len + 1 */
  const int64_t* input_indices; /* n_terms, upper-triangle input index */
  const int64_t* coeff_indices; /* n_terms, -1 = static (coefficient 1) */
  const uint8_t* coeff_conj;
  const uint8_t* input_conj;
  const double* factors_re;
  const double* factors_im;
  /* Differentiable plans (lb_adjoint): n_diff_params base slots p_m.  The
   * expression list then holds, after the n_coefficients coefficient
   * programs, n_coefficients * n_diff_params derivative programs
   * (k * n_diff_params + m is d c_k / d p_m); compounds may include
   * derivative compounds they read.  0 for an ordinary plan. */
  int32_t n_diff_params;
  const int32_t* diff_param_slots;
  /* 1: the derivative programs carry one more column, d c_k / d t (the
   * programs are then K * (n_diff_params + 1), k * (n_diff_params + 1) +
   * n_diff_params being the time derivative), which lb_adjoint needs to
   * differentiate the step-size controller (LB_ADJOINT_ADAPTIVE).  0 for an
   * ordinary or frozen-step plan. */
  int32_t diff_time;
} lb_plan_desc;

typedef struct lb_plan lb_plan;

/*
 * Build a device plan: validates the description (same checks as
 * rust/src/lindblad/plan.rs:1069-1153), classifies coefficients as
 * time-dependent (plan.rs:1235-1291), assigns upper outputs to lanes and
 * uploads the tables to the current HIP device.
 * Replaces prepare_lindblad_problem_py (python_api.rs:271-274).
 */
int lb_plan_create(const lb_plan_desc* desc, lb_plan** out_plan);
void lb_plan_destroy(lb_plan* plan);

/*
 * Plan-specialised paired kernels (no counterpart in the reference, whose
 * Rust RHS interprets the coefficient programs on every call,
 * rust/src/lindblad/eval.rs:791-911).  With enable != 0, the plan's
 * time-dependent compound and coefficient programs are emitted as
 * straight-line device code; the first lb_solve call that takes a paired
 * variant for a solver compiles that variant with the code at run time
 * (hiprtc, about 2 s) and later calls reuse it.  Values are bitwise those of
 * the expression machine.  A plan without time dependence is left as it is.
 * Returns LB_ERR_UNSUPPORTED (with the reason) when an instruction has no
 * generated form.  When the run-time compilation or the module load of a
 * variant fails, lb_solve runs the precompiled kernel instead (bitwise the
 * same results), never tries that variant again for the plan, and records
 * the reason for lb_plan_specialize_status.  The environment variable
 * CENTREX_LB_RTC_OPTIONS, when set, is appended to the compile options
 * (whitespace-separated; a diagnostic hook).
 * enable = 0 returns the plan to the precompiled kernels.
 */
int lb_plan_specialize(lb_plan* plan, int32_t enable);

/*
 * Whether a specialised compilation of this plan has failed: returns 0 when
 * none has, 1 when one has, with its reason copied (NUL-terminated,
 * truncated to cap bytes) into buf when buf is not NULL.  No reference
 * counterpart.
 */
int lb_plan_specialize_status(const lb_plan* plan, char* buf, int64_t cap);

/* Thread-local message for the last failing call on this thread. */
const char* lb_last_error(void);

/*
 * Plan facts, written to info[0..15]:
 *   0 n_states, 1 dim, 2 upper_len, 3 n_terms, 4 n_coefficients,
 *   5 time_dependent (0/1), 6 n_published (elements shared through LDS),
 *   7 n_gather_terms, 8 n_self_static_terms, 9 slot_count (padded),
 *   10 gather_rows (sum over 64-lane blocks of the block's max gather count),
 *   11 flops_per_rhs (algorithmic, see lb_solve_args.out_stats),
 *   12 dynamic coefficient count, 13 lds bytes per trajectory.
 */
int lb_plan_info(const lb_plan* plan, int64_t* info);

/*
 * The same facts for the plan as lb_solve runs it with initial support
 * `y0_support` (host, dim bytes, as in lb_solve_args): 1 becomes the number
 * of packed reals integrated, 2 the active upper elements, 3 their term
 * count, 6-11 and 13-15 those of the restricted tables.  NULL gives the full
 * plan.  No reference counterpart.
 */
int lb_plan_support_info(const lb_plan* plan, const uint8_t* y0_support, int64_t* info);

/*
 * Batched RHS evaluation: out[b] = d(rho)/dt at (y[b], t[b]) with per-batch
 * overrides of base parameter slots.  Replaces lindblad_rhs_py /
 * rhs_packed_py (python_api.rs:276-352) for a batch; used to check the RHS
 * kernel against the oracle.
 *   y, out: device, batch x dim;  t: device, batch;
 *   override_slots: host, n_override base-slot indices;
 *   override_values: device, batch x n_override complex (re, im interleaved).
 */
int lb_rhs_eval(const lb_plan* plan, int64_t batch, const double* y, const double* t,
                int32_t n_override, const int32_t* override_slots,
                const double* override_values, double* out, void* stream);

/*
 * Arguments of lb_solve.  Trajectory b integrates y0 (row b of y0, or the
 * shared row when y0_shared) from t0 to t1 (t1_per_traj[b] when given) with
 * base-slot overrides taken either from override_values (row b) or, for a
 * grid, from the Cartesian product of the axes with the last axis fastest
 * (rust/src/lindblad/ode_batch.rs:353-391).
 *
 * Integral output follows the reference's trapezoid over pushed points
 * (rust/src/ode/output.rs:315-330): t0 (when save_start) and every accepted
 * step.  Populations / selected / full are the state at t1.
 *
 * out_stats (device, n_traj x 4 int64): accepted, rejected, rhs_calls,
 * status (LB_TRAJ_*); out_fail_t (device, n_traj): x at failure.
 */
typedef struct lb_solve_args {
  int64_t n_traj;
  int32_t solver; /* LB_SOLVER_* */
  double t0;
  double t1;
  const double* t1_per_traj; /* device or NULL */
  double abstol;
  double reltol;
  double dt;
  int64_t maxiters;
  int32_t save_start;
  const double* y0; /* device */
  int32_t y0_shared;
  int32_t n_override;
  const int32_t* override_slots;  /* host */
  const double* override_values;  /* device, n_traj x n_override complex; NULL for grid */
  int32_t n_axes;                 /* grid form when > 0 (n_axes == n_override) */
  const int64_t* axis_lengths;    /* host, n_axes */
  const double* axis_values;      /* device, sum(axis_lengths) complex */
  int32_t output_kind;            /* LB_OUT_* */
  int32_t n_selected;
  const int32_t* selected_rows;   /* host, n_selected (full-matrix indices) */
  const int32_t* selected_cols;
  int32_t n_weights;
  const int32_t* weight_indices;  /* host, n_weights diagonal indices */
  const double* weight_values;    /* host */
  double* out_values;             /* device, n_traj x width (selected: complex pairs) */
  int64_t* out_stats;             /* device, n_traj x 4 */
  double* out_fail_t;             /* device, n_traj */
  int32_t mode;                   /* LB_MODE_* */
  int32_t threads_per_traj;       /* 0 = plan default; 64..1024, multiple of 64 */
  int32_t workgroups;             /* persistent grid size, 0 = fill the device */
  double* out_integral;           /* device, n_traj, or NULL: the weighted trapezoid
                                     alongside a non-integral output (one pass
                                     instead of the reference's two solves) */
  int32_t variant;                /* kernel variant: 0 = plan default, k = variant k-1
                                     (slots per thread / occupancy; results are
                                     bitwise independent of it) */
  int32_t n_saveat;               /* > 0: push points at these times (dense output inside steps),
                                     otherwise at t0 (save_start) and every accepted step */
  const double* saveat;           /* device, ascending, within [t0, t1] (the save plan of
                                     rust/src/ode/common.rs:124-151, t0 included when saving
                                     the start) */
  int32_t trace;                  /* 1: out_values holds one row per pushed point,
                                     n_traj x capacity x width (output_when="saveat") */
  int64_t capacity;               /* rows per trajectory in trace mode */
  double* out_times;              /* device, n_traj x capacity times of the rows, or NULL */
  int32_t* out_counts;            /* device, n_traj: rows pushed (trace mode) */
  const uint8_t* y0_support;      /* host, dim bytes, or NULL: nonzero where any
                                     trajectory's y0 may be nonzero.  The solve then
                                     runs only the elements reachable from that
                                     support through the RHS program (all others
                                     stay exactly zero) and reports
                                     LB_TRAJ_BAD_SUPPORT for a y0 outside it. */
  /* Adjoint tape: lb_solve records it when tape_h is not NULL; lb_adjoint
   * reads it.  One record per step ATTEMPT (accepted or rejected), in order:
   * tape_t its start time, tape_h its step size (after the last-step clip),
   * tape_c (err, h_next, fac_old): the error norm, the step size the
   * controller proposed next (before the h_max and after-reject rules) and
   * the controller's fac_old when the attempt started (common.rs:48-69),
   * tape_f its LB_TAPE_* flags; tape_y holds the state at the start of every
   * ckpt_every-th attempt.  All device arrays: tape_t, tape_h, tape_f
   * n_traj x tape_cap; tape_c n_traj x tape_cap x 3; tape_y n_traj x
   * ceil(tape_cap / ckpt_every) x slots x 2 (lb_plan_support_info field 9
   * gives the slots); tape_n n_traj (attempts recorded).  More attempts than
   * tape_cap: LB_TRAJ_CAPACITY. */
  double* tape_t;
  double* tape_h;
  double* tape_y;
  int32_t* tape_n;
  int64_t tape_cap;
  int32_t ckpt_every;
  double* out_grad;               /* lb_adjoint: device, n_traj x n_diff_params */
  const double* replay_h;         /* device, n_traj x tape_cap, or NULL: integrate with these
                                     accepted step sizes (a recorded tape) instead of the
                                     step-size controller, for checks of the adjoint */
  const int32_t* replay_n;        /* device, n_traj: steps of each schedule */
  const int32_t* queue_order;     /* device, n_traj, or NULL: the work queue hands out
                                     trajectories in this order (a permutation of
                                     0..n_traj-1; e.g. longest first for uneven batches).
                                     Entries outside [0, n_traj) are skipped. */
  /* Terminal stop event (SolverOptions stop_event, rust/src/lindblad/ode_impl.rs:8-16).
   * The event function g(t, y) is checked at the end of every accepted step; a
   * sign change or zero since the previous accepted step ends the trajectory at
   * the root: 60 bisections on the dense output for the adaptive solvers
   * (dopri5.rs:280-317, tsit5.rs:274-309), the step end for the fixed-step ones
   * (fixed.rs:213-238).  The root is pushed once as the last point; save-plan
   * points past it are not pushed and are not an error.  Final outputs hold the
   * state at the root.  Runs on the multi-wave kernels (the single-wave and
   * paired forms are not used for such a call); no tapes or replays. */
  int32_t event_kind;             /* LB_EVENT_* */
  int32_t n_event_indices;        /* population: diagonal indices summed in this order */
  const int32_t* event_indices;   /* host */
  double event_threshold;
  int32_t n_event_instr;          /* expression: instructions in the lb_plan_desc format */
  const int32_t* event_op;        /* host, n_event_instr each */
  const int32_t* event_index;
  const int32_t* event_argc;
  const int32_t* event_function;
  const double* event_re;
  const double* event_im;
  double* out_event_t;            /* device, n_traj, or NULL: the root's time, NaN when the
                                     event did not fire */
  /* the rest of the attempt tape (see tape_t above) */
  double* tape_c;
  int32_t* tape_f;
  int32_t adjoint_mode;           /* lb_adjoint: LB_ADJOINT_ADAPTIVE or LB_ADJOINT_FROZEN */
} lb_solve_args;

/* Flags of an attempt-tape record (lb_solve_args.tape_f). */
#define LB_TAPE_ACCEPTED 1 /* the attempt was accepted (err <= 1)                        */
#define LB_TAPE_CLIPPED 2  /* h was clipped to t1 - t (dopri5.rs:232-235)                 */
#define LB_TAPE_CLAMPED 4  /* the controller's factor hit facc1 or facc2 (no err/fac_old dependence) */
#define LB_TAPE_HMAX 8     /* accepted: the next step was cut to h_max                   */
#define LB_TAPE_MINRULE 16 /* accepted after a reject: the next step is |h| (common.rs:59-61) */
#define LB_TAPE_ERRQ 32    /* accepted with err > 1e-4: fac_old takes err                */
#define LB_TAPE_FIRST 64   /* the first accepted step                                    */

/* lb_adjoint modes. */
#define LB_ADJOINT_ADAPTIVE 0 /* the derivative of the adaptive solve: the accepted and rejected
                                 attempts, the error norm and the step-size controller are
                                 differentiated (needs a plan with diff_time = 1) */
#define LB_ADJOINT_FROZEN 1   /* the accepted step sizes held fixed: the derivative of a replay */

/*
 * Integrate a batch of trajectories.  Replaces solve_lindblad_batch_ode_py
 * (python_api.rs:530-669) and solve_lindblad_grid_ode_py (python_api.rs:671-799)
 * for output_when="final" and the accepted-step integral; the single-trajectory
 * solve_lindblad_ode_py (python_api.rs:422-528) is the n_traj = 1 case.
 * Stream-ordered on `stream` (a hipStream_t, NULL = default stream).
 */
int lb_solve(const lb_plan* plan, const lb_solve_args* args, void* stream);

/*
 * Discrete adjoint of a recorded solve (no reference counterpart: the
 * reference has no gradients).  For the weighted trapezoid integral of each
 * trajectory (output_kind LB_OUT_INTEGRAL, pushes at t0 when save_start and
 * at every accepted step, no saveat), writes out_grad[b, m] = d I_b / d p_m
 * for the plan's differentiated parameters (lb_plan_desc.n_diff_params).
 * With adjoint_mode LB_ADJOINT_ADAPTIVE that is the derivative of the
 * adaptive solve itself: every attempt of the tape, the error norm
 * (dopri5.rs:88-102, tsit5.rs:104-119), the controller (common.rs:48-69),
 * the last-step clip and the trapezoid's step lengths (output.rs:315-330)
 * are differentiated.  LB_ADJOINT_FROZEN holds the accepted step sizes
 * fixed.  `args` must describe the same trajectories, parameters, y0
 * support and tape as the lb_solve that recorded the tape.  The library
 * keeps a per-plan workspace of ckpt_every recomputed attempts per resident
 * workgroup.  A trajectory whose tape is empty or longer than tape_cap gets
 * NaN gradients; so does every trajectory a workgroup of the single-block
 * adjoint handles after one of its bounded waits between waves ran out.
 */
int lb_adjoint(const lb_plan* plan, const lb_solve_args* args, void* stream);

/*
 * lb_adjoint with no device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  The work-queue counter and
 * the adjoint's workspace live in a fixed device array of
 * LB_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the launch
 * uses as many persistent workgroups as fit in it (fewer than lb_adjoint may
 * use, with the same results).  The plan's tables for args->y0_support must
 * exist already (the lb_solve that recorded the tape builds them).  Calls
 * that overlap on one GPU share the array and race.  LB_ERR_UNSUPPORTED when
 * one workgroup's workspace (ckpt_every attempts) exceeds the array.
 */
#define LB_NOALLOC_SCRATCH_BYTES (64u << 20)
int lb_adjoint_noalloc(const lb_plan* plan, const lb_solve_args* args, void* stream);

/*
 * Several plans in one call (no reference counterpart: the reference solves
 * one plan per call, batch.py:209).  Group g integrates args[g] with
 * plans[g] as lb_solve would; the groups run concurrently on library-owned
 * streams forked from `stream` and joined back into it, so a HIP graph
 * captured on `stream` holds the whole uneven batch as one launch.  Each
 * group's args.workgroups bounds its share of the device (0 lets a group fill
 * the device, which serialises the groups).  At most LB_MAX_GROUPS groups.
 */
#define LB_MAX_GROUPS 16

/*
 * The launch a persistent lb_solve of this plan would make for the given y0
 * support (NULL = every element), variant and threads_per_traj (0 = plan
 * defaults): out[0] variant (1-based), out[1] threads per workgroup,
 * out[2] resident workgroups per CU, out[3] CUs of the device.  No reference
 * counterpart: it sizes the groups of lb_solve_group.
 */
int lb_launch_info(const lb_plan* plan, const uint8_t* y0_support, int32_t variant,
                   int32_t threads_per_traj, int32_t* out);
int lb_solve_group(const lb_plan* const* plans, const lb_solve_args* args, int32_t n_groups,
                   void* stream);

/*
 * Kernel variant `id` (0-based): slots per thread, largest workgroup and the
 * waves-per-SIMD occupancy hint it was compiled for.  Returns LB_ERR_INVALID
 * past the last variant.  No reference counterpart: it feeds the autotuner.
 */
int lb_variant_info(int32_t id, int32_t* slots_per_thread, int32_t* max_threads,
                    int32_t* waves_per_simd);

/* ------------------------------------------------------------------------
 * Effective-Hamiltonian Lindblad plans (rust/src/effective_lindblad/).
 *
 * The state is the split-real density matrix y = [Re rho.flat, Im rho.flat]
 * (real_dim = 2 n^2 doubles, centrex_tlf/effective_hamiltonian/rust_plan.py:18-39)
 * and the right-hand side (rust/src/effective_lindblad/rhs.rs:86-138) is
 *
 *   dy = A_comb(x) y + (Omega / 2) A_opt(x) y + delta A_det(x) y
 *
 * with x, Omega and delta the real parts of three parameter slots at t.  Each
 * operator is CSR over real_dim rows; every stored entry is a cubic in
 * dx = x - field_grid[k] on grid interval k, coefficients c0..c3 at
 * coeffs[(k * nnz + entry) * 4 ...] (rust/src/effective_lindblad/plan.rs:18-40).
 * x outside the grid by more than 1e-10 of its span fails the trajectory
 * (LB_TRAJ_FIELD_BELOW / LB_TRAJ_FIELD_ABOVE, fail_t = x).
 * ------------------------------------------------------------------------ */
#define LB_TRAJ_FIELD_BELOW 10 /* "field coordinate {x:.6e} is below the operator grid minimum {g:.6e}" */
#define LB_TRAJ_FIELD_ABOVE 11 /* "field coordinate {x:.6e} is above the operator grid maximum {g:.6e}" */

#define LB_EFF_OUT_POPULATIONS 0 /* y[i*n + i], width n */
#define LB_EFF_OUT_SELECTED 1    /* (re, im) of rho_ij, width 2 * n_selected */
#define LB_EFF_OUT_INTEGRAL 2    /* trapezoid of sum_k w_k y[idx_k] over the pushed points, width 1 */
#define LB_EFF_OUT_FULL 3        /* y, width real_dim */

typedef struct lb_eff_plan_desc {
  int32_t n_states;
  int32_t real_dim;               /* 2 * n_states^2 */
  int32_t n_grid;                 /* >= 2 */
  const double* field_grid;       /* n_grid, ascending */
  const int64_t* row_ptrs[3];     /* combined, optical, detuning: real_dim + 1 each */
  const int64_t* col_indices[3];  /* nnz[k] each */
  const double* coeffs[3];        /* (n_grid - 1) * nnz[k] * 4 each */
  int64_t nnz[3];
  int32_t field_slot, rabi_slot, detuning_slot;
  /* parameter graph, in the lb_plan_desc format (no coefficients) */
  int32_t n_slots;
  int32_t n_base;
  const double* base_re;
  const double* base_im;
  int32_t n_compounds;
  const int32_t* compound_slots;
  const int32_t* expr_ptrs;       /* n_compounds + 1 */
  const int32_t* instr_op;
  const int32_t* instr_index;
  const int32_t* instr_argc;
  const int32_t* instr_function;
  const double* instr_re;
  const double* instr_im;
} lb_eff_plan_desc;

typedef struct lb_eff_plan lb_eff_plan;

/* Replaces prepare_effective_lindblad_plan_py (rust/src/effective_lindblad/python_api.rs:16-26). */
int lb_eff_plan_create(const lb_eff_plan_desc* desc, lb_eff_plan** out_plan);
void lb_eff_plan_destroy(lb_eff_plan* plan);

/*
 * Arguments of lb_eff_solve: trajectory b integrates the shared y0 from t0 to
 * t1 with base-slot overrides override_values[b * n_override + m] (real, as
 * rust/src/effective_lindblad/rhs.rs:26-44).  Points are pushed at t0 (when
 * save_start) and every accepted step, or at the save plan `saveat` (device,
 * ascending, t0 first when saving the start) through the dense output; each
 * push writes a row of out_values (n_traj x capacity x width) and out_times,
 * except for the integral, which writes its final value (one row).
 */
typedef struct lb_eff_solve_args {
  int64_t n_traj;
  int32_t solver;                 /* LB_SOLVER_DOPRI5 or LB_SOLVER_TSIT5 */
  double t0, t1, abstol, reltol, dt;
  int64_t maxiters;
  int32_t save_start;
  const double* y0;               /* device, real_dim */
  int32_t n_override;
  const int32_t* override_slots;  /* host */
  const double* override_values;  /* device, n_traj x n_override */
  int32_t output_kind;            /* LB_EFF_OUT_* */
  int32_t n_selected;
  const int32_t* selected_rows;   /* host */
  const int32_t* selected_cols;
  int32_t n_weights;
  const int32_t* weight_indices;  /* host, indices into y */
  const double* weight_values;    /* host */
  int32_t n_saveat;
  const double* saveat;           /* device */
  int64_t capacity;               /* rows per trajectory */
  double* out_values;             /* device */
  double* out_times;              /* device, n_traj x capacity, or NULL */
  int32_t* out_counts;            /* device, n_traj: points pushed */
  int64_t* out_stats;             /* device, n_traj x 4 as lb_solve_args */
  double* out_fail_t;             /* device, n_traj */
  int32_t workgroups;             /* 0 = fill the device */
} lb_eff_solve_args;

/* Replaces solve_effective_lindblad_py / solve_effective_lindblad_batch_py
 * (rust/src/effective_lindblad/python_api.rs:28-91, 152-250). */
int lb_eff_solve(const lb_eff_plan* plan, const lb_eff_solve_args* args, void* stream);

/* Batched RHS: out[b] = dy at (y[b], t[b]) (device arrays, batch rows); the
 * counterpart of debug_effective_lindblad_rhs_py (python_api.rs:93-150).
 * out_status[b] (device) gets 0 or an LB_TRAJ_FIELD_* code. */
int lb_eff_rhs_eval(const lb_eff_plan* plan, int64_t batch, const double* y, const double* t,
                    double* out, int32_t* out_status, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CENTREX_LINDBLAD_H */
