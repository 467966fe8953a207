/*
 * CPU oracle of the Lindblad hot path: a plain-C restatement of the
 * reference's Rust solver, used ONLY as the checker in tests, in
 * __graft_entry__.smoke() and as bench.py's cpu_baseline leg.  Nothing in the
 * product path links or calls it.
 */
#ifndef LB_ORACLE_H
#define LB_ORACLE_H

#include <stdint.h>

#include "../include/centrex_lindblad.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbo_plan lbo_plan;

#define LBO_OUT_POPULATIONS 0
#define LBO_OUT_SELECTED 1
#define LBO_OUT_INTEGRAL 2
#define LBO_OUT_FULL 3
#define LBO_OUT_RATE 4

typedef struct lbo_solve_args {
  int32_t solver; /* LB_SOLVER_DOPRI5 / LB_SOLVER_TSIT5 */
  double t0;
  double t1;
  double abstol;
  double reltol;
  double dt;
  int64_t maxiters;
  int32_t save_start;
  const double* saveat; /* NULL = push every accepted step */
  int64_t n_saveat;
  int32_t output_kind;
  int32_t store_trace; /* integral: keep the running value at every push */
  int32_t n_selected;
  const int32_t* selected_rows;
  const int32_t* selected_cols;
  int32_t n_weights;
  const int32_t* weight_indices;
  const double* weight_values;
  /* terminal stop event (LB_EVENT_*; rust/src/lindblad/ode_impl.rs:134-153):
   * population indices + threshold, or an expression in the lb_plan_desc
   * instruction format; event_t (n_traj, or NULL) receives each root's time,
   * NaN where the event did not fire */
  int32_t event_kind;
  int32_t n_event_indices;
  const int32_t* event_indices;
  double event_threshold;
  int32_t n_event_instr;
  const int32_t* event_op;
  const int32_t* event_index;
  const int32_t* event_argc;
  const int32_t* event_function;
  const double* event_re;
  const double* event_im;
  double* event_t;
  /* n_traj, or NULL: the weighted trapezoid over t0 (save_start) and every
   * accepted step, computed alongside the output (the integral solve the
   * reference would run separately with saveat None; adaptive solvers, no
   * stop event) */
  double* integral_out;
  int32_t integral_save_start; /* push t0 into integral_out's trapezoid */
  int32_t saveat_is_t1;        /* the save plan is {t1} of each trajectory
                                  (output_when="final" with per-trajectory end times) */
} lbo_solve_args;

int lbo_plan_create(const lb_plan_desc* desc, lbo_plan** out);
/* An effective-Hamiltonian plan: the same solvers and batch driver over the
 * split-real state and the interpolated-operator RHS. */
int lbo_eff_plan_create(const lb_eff_plan_desc* desc, lbo_plan** out);
void lbo_plan_destroy(lbo_plan* plan);
const char* lbo_last_error(void);
int64_t lbo_plan_terms(const lbo_plan* plan);

/* One RHS evaluation; values are n_override complex pairs. */
int lbo_rhs(lbo_plan* plan, const double* y, double t, int32_t n_override,
            const int32_t* override_slots, const double* override_values, double* out);

/* The starting step of hinit (rust/src/ode/common.rs:72-118) for y at t0
 * towards t1, with the base-slot overrides of lbo_rhs. */
int lbo_hinit(lbo_plan* plan, const double* y, double t0, double t1, double abstol, double reltol,
              int32_t n_override, const int32_t* override_slots, const double* override_values,
              double* h_out);

/* One expression program in the lb_plan_desc instruction format over
 * `slots` (n_slots complex pairs) at time t; out = (re, im). */
int lbo_eval_expression(int32_t n_instr, const int32_t* op, const int32_t* index, const int32_t* argc,
                        const int32_t* function, const double* re, const double* im, int32_t n_slots,
                        const double* slots, double t, double* out);

/* Output width in doubles per pushed point for these arguments. */
int64_t lbo_output_width(const lbo_plan* plan, const lbo_solve_args* args);

/*
 * Batch of trajectories on `threads` worker threads (0 = all cores) with
 * dynamic chunking.  Per trajectory b: times[b*capacity ...], values[b*capacity*width ...],
 * n_times[b], stats[b*4 + {0 accepted, 1 rejected, 2 rhs_calls, 3 status}],
 * fail_x[b].  status uses LB_TRAJ_* plus 5 = other error (message via
 * lbo_last_error for the first failing trajectory).
 */
int lbo_solve_batch(const lbo_plan* plan, const lbo_solve_args* args, int64_t n_traj,
                    const double* y0, int32_t y0_shared, const double* t1_per_traj,
                    int32_t n_override, const int32_t* override_slots,
                    const double* override_values, int32_t threads, int64_t capacity,
                    double* times, double* values, int64_t* n_times, int64_t* stats,
                    double* fail_x);

#ifdef __cplusplus
}
#endif

#endif
