/*
 * CPU oracle of the Lindblad hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's Rust solver, used as the checker by
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The
 * product path (the HIP library under centrex-tlf_amd/) never links or calls
 * it.  Each part follows the reference file it restates:
 *
 *   complex arithmetic      num-complex 0.4.6 formulas used by eval.rs:135-153
 *   expression machine      rust/src/lindblad/eval.rs:155-451, 664-911
 *   parameter graph         rust/src/lindblad/plan.rs:118-174
 *   dynamic/time flags      rust/src/lindblad/plan.rs:1235-1291
 *   upper -> packed map     rust/src/lindblad/rhs.rs:151-176
 *   partitioned kernel      rust/src/lindblad/rhs.rs:285-384, 815-940, 1789-1867  (n <= 40)
 *   split-input kernel      rust/src/lindblad/rhs.rs:207-250, 633-695, 786-813, 1869-1968 (n > 40)
 *   controller, hinit, save plan   rust/src/ode/common.rs:19-151
 *   Dopri5                  rust/src/ode/dopri5.rs:5-379
 *   Tsit5                   rust/src/ode/tsit5.rs:5-363
 *   collectors              rust/src/ode/output.rs:70-384
 *   batch driver            rust/src/lindblad/ode_batch.rs:176-351 (rayon map_init -> pthreads)
 *
 * Build with -ffp-contract=off so that a*b + c stays two roundings, as in the
 * Rust code (no implicit FMA).  Terminal stop events are not restated.
 *
 * Parity status of each part is recorded in tests/test_oracle_golden.py; the
 * ULP-level behaviour of the transcendental builtins (libm vs Rust std) is
 * "parity unpinned": it is pinned only at the 1e-7 result level through the
 * reference-held golden vectors.
 */
#define _GNU_SOURCE
#include "lb_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef struct {
  double re, im;
} cpx;

static __thread char g_err[512];

static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return LB_ERR_INVALID;
}

const char* lbo_last_error(void) { return g_err; }

/* ---------------- complex arithmetic (num-complex formulas) -------------- */

static inline cpx C(double re, double im) {
  cpx z = {re, im};
  return z;
}
static inline cpx cadd(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
static inline cpx csub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
static inline cpx cmul(cpx a, cpx b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cpx cdiv(cpx a, cpx b) {
  double ns = b.re * b.re + b.im * b.im;
  double re = a.re * b.re + a.im * b.im;
  double im = a.im * b.re - a.re * b.im;
  return C(re / ns, im / ns);
}
static inline cpx cconj(cpx a) { return C(a.re, -a.im); }
static inline cpx cneg(cpx a) { return C(-a.re, -a.im); }
static inline cpx from_polar(double r, double th) { return C(r * cos(th), r * sin(th)); }
static cpx cexp_nc(cpx z) { return from_polar(exp(z.re), z.im); }
static cpx csin_nc(cpx z) { return C(sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)); }
static cpx ccos_nc(cpx z) { return C(cos(z.re) * cosh(z.im), -(sin(z.re) * sinh(z.im))); }
static cpx ctan_nc(cpx z) {
  double tr = z.re + z.re, ti = z.im + z.im;
  double d = cos(tr) + cosh(ti);
  return C(sin(tr) / d, sinh(ti) / d);
}
/* x^y = exp(y ln x) in polar form; x^0 = 1, 0^y = 0. */
static cpx cpow_nc(cpx a, cpx b) {
  if (b.re == 0.0 && b.im == 0.0) return C(1.0, 0.0);
  double r = hypot(a.re, a.im);
  if (r == 0.0) return C(0.0, 0.0);
  double th = atan2(a.im, a.re);
  return from_polar(pow(r, b.re) * exp(-b.im * th), b.re * th + b.im * log(r));
}

/* ---------------- expression machine ------------------------------------ */

enum {
  OP_CONST = 1, OP_SLOT, OP_TEMP, OP_TIME, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_NEG,
  OP_CONJ, OP_BUILTIN, OP_HELPER, OP_GT, OP_GE, OP_LT, OP_LE, OP_EQ, OP_NE
};

typedef struct {
  int32_t op, index, argc, function;
  double re, im;
} instr_t;

typedef struct {
  int32_t start, end;
  int uses_time;
} expr_t;

#define STACK_MAX 1024
static const double ED_XTB = 0.315 * 1.602176634e-19 * 5.29177210544e-11;

static double rem_euclid(double x, double m) {
  double r = fmod(x, m);
  return r < 0.0 ? r + fabs(m) : r;
}
static double gaussian_2d(double x, double y, double a, double mx, double my, double sx, double sy) {
  double dx = x - mx, dy = y - my;
  return a * exp(-((dx * dx) / (2.0 * sx * sx) + (dy * dy) / (2.0 * sy * sy)));
}
static double rabi_from_intensity(double intensity, double coupling, double dipole) {
  double hbar = 1.0545718176461565e-34, c = 299792458.0, eps0 = 8.8541878128e-12;
  double e = sqrt(intensity * 2.0 / (c * eps0));
  return e * coupling * dipole / hbar;
}

/* as_real (eval.rs:95-101): a helper's real argument. */
static int as_real(cpx v, double* out) {
  if (fabs(v.im) > 1e-12) return fail("complex value used where real scalar was required");
  *out = v.re;
  return 0;
}

/* multipass_2d_intensity (eval.rs:250-275): a = x, y, amplitudes[L],
 * xlocs[L], ylocs[L], sigma_x, sigma_y. */
static int multipass(const cpx* a, int L, double* out) {
  double x, y, sx, sy, intensity = 0.0;
  int rc = as_real(a[0], &x) || as_real(a[1], &y) || as_real(a[2 + 3 * L], &sx) || as_real(a[3 + 3 * L], &sy);
  if (rc) return LB_ERR_INVALID;
  for (int k = 0; k < L; ++k) {
    double amp, xl, yl;
    if (as_real(a[2 + k], &amp) || as_real(a[2 + L + k], &xl) || as_real(a[2 + 2 * L + k], &yl)) return LB_ERR_INVALID;
    intensity += gaussian_2d(x, y, amp, xl, yl, sx, sy);
  }
  *out = intensity;
  return 0;
}

/* linear_interp (eval.rs:360-388): a = x, grid[n], values[n]. */
static int linear_interp(const cpx* a, int n, double* out) {
  double x, first;
  const cpx* grid = a + 1;
  const cpx* values = a + 1 + n;
  if (as_real(a[0], &x)) return LB_ERR_INVALID;
  if (n == 1) return as_real(values[0], out);
  if (as_real(grid[0], &first)) return LB_ERR_INVALID;
  if (x <= first) return as_real(values[0], out);
  for (int idx = 0; idx + 1 < n; ++idx) {
    double left, right;
    if (as_real(grid[idx], &left) || as_real(grid[idx + 1], &right)) return LB_ERR_INVALID;
    if (right <= left) return fail("linear_interp grid must be strictly increasing");
    if (x <= right) {
      double yl, yr;
      if (as_real(values[idx], &yl) || as_real(values[idx + 1], &yr)) return LB_ERR_INVALID;
      double frac = (x - left) / (right - left);
      *out = yl + frac * (yr - yl);
      return 0;
    }
  }
  return as_real(values[n - 1], out);
}

/* Slopes of a monotone cubic (pchip) on knots g, v (eval.rs:318-331,
 * plan.rs:54-66); h and delta are the interval widths and secants. */
static int pchip_slopes(const double* g, const double* v, int n, double* h, double* delta, double* d) {
  for (int i = 0; i + 1 < n; ++i) {
    h[i] = g[i + 1] - g[i];
    if (h[i] <= 0.0) return fail("pchip_interp grid must be strictly increasing");
    delta[i] = (v[i + 1] - v[i]) / h[i];
  }
  d[0] = delta[0];
  d[n - 1] = delta[n - 2];
  for (int i = 1; i + 1 < n; ++i) {
    double s0 = delta[i - 1] < 0.0 ? -1.0 : 1.0, s1 = delta[i] < 0.0 ? -1.0 : 1.0;
    if (s0 != s1 || delta[i - 1] == 0.0 || delta[i] == 0.0) {
      d[i] = 0.0;
    } else {
      double w1 = 2.0 * h[i] + h[i - 1], w2 = h[i] + 2.0 * h[i - 1];
      d[i] = (w1 + w2) / (w1 / delta[i - 1] + w2 / delta[i]);
    }
  }
  return 0;
}

/* pchip_interp, Hermite form (eval.rs:290-351) when table == 0; through a
 * precomputed table in Horner form (plan.rs:33-115) when table != 0.
 * a = x, grid[n], values[n]. */
static int pchip(const cpx* a, int n, int table, double* out) {
  double x;
  if (as_real(a[0], &x)) return LB_ERR_INVALID;
  if (n < 2) return as_real(a[2], out);
  double* buf = (double*)malloc(sizeof(double) * 5 * (size_t)n);
  if (!buf) return fail("out of memory");
  double *g = buf, *v = buf + n, *h = buf + 2 * n, *delta = buf + 3 * n, *d = buf + 4 * n;
  for (int i = 0; i < n; ++i) {
    g[i] = a[1 + i].re;
    v[i] = a[1 + n + i].re;
  }
  int rc = pchip_slopes(g, v, n, h, delta, d);
  if (rc == 0 && !table) {
    if (x <= g[0]) {
      *out = v[0];
    } else if (x >= g[n - 1]) {
      *out = v[n - 1];
    } else {
      int idx = 0;
      for (int i = 0; i + 1 < n; ++i)
        if (x < g[i + 1] || i == n - 2) { idx = i; break; }
      double t = (x - g[idx]) / h[idx], t2 = t * t, t3 = t2 * t;
      double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
      double h01 = -2.0 * t3 + 3.0 * t2, h11 = t3 - t2;
      *out = h00 * v[idx] + h10 * h[idx] * d[idx] + h01 * v[idx + 1] + h11 * h[idx] * d[idx + 1];
    }
  } else if (rc == 0) {
    int i;
    double dx;
    if (x <= g[0]) {
      i = 0;
      dx = 0.0;
    } else if (x >= g[n - 1]) {
      i = n - 2;
      dx = g[n - 1] - g[n - 2];
    } else {
      int pp = 0; /* partition_point(g <= x) */
      while (pp < n && g[pp] <= x) ++pp;
      i = pp - 1 < n - 2 ? pp - 1 : n - 2;
      dx = x - g[i];
    }
    double c0 = v[i], c1 = d[i];
    double c2 = (3.0 * delta[i] - 2.0 * d[i] - d[i + 1]) / h[i];
    double c3 = (d[i] + d[i + 1] - 2.0 * delta[i]) / (h[i] * h[i]);
    *out = (x <= g[0]) ? c0 : c0 + dx * (c1 + dx * (c2 + dx * c3));
  }
  free(buf);
  return rc;
}

static int apply_helper(int id, const cpx* a, int argc, cpx* out) {
#define R(i) (a[i].re)
  double r = 0.0;
  switch (id) {
    case 8:
      if (argc < 4 || (argc - 4) % 3) return fail("multipass_2d_intensity with %d arguments", argc);
      if (multipass(a, (argc - 4) / 3, &r)) return LB_ERR_INVALID;
      *out = C(r, 0.0);
      return 0;
    case 10: {
      if (argc < 5 || ((argc - 5) % 3 && (argc - 6) % 3)) return fail("multipass_2d_rabi with %d arguments", argc);
      int L = (argc - 5) % 3 == 0 ? (argc - 5) / 3 : (argc - 6) / 3;
      double coupling, dipole = ED_XTB;
      if (multipass(a, L, &r) || as_real(a[4 + 3 * L], &coupling)) return LB_ERR_INVALID;
      if (argc == 6 + 3 * L && as_real(a[5 + 3 * L], &dipole)) return LB_ERR_INVALID;
      *out = C(rabi_from_intensity(r, coupling, dipole), 0.0);
      return 0;
    }
    case 14:
      if (argc < 3 || argc % 2 == 0) return fail("linear_interp with %d arguments", argc);
      if (linear_interp(a, (argc - 1) / 2, &r)) return LB_ERR_INVALID;
      *out = C(r, 0.0);
      return 0;
    case 16:
    case 17:
      if (argc < 3 || argc % 2 == 0) return fail("pchip_interp with %d arguments", argc);
      if (pchip(a, (argc - 1) / 2, id == 17, &r)) return LB_ERR_INVALID;
      *out = C(r, 0.0);
      return 0;
    case 1:
      *out = C(gaussian_2d(R(0), R(1), R(2), R(3), R(4), R(5), R(6)), 0.0);
      return 0;
    case 2: {
      double th = R(7), ct = cos(th), st = sin(th), sx = R(5), sy = R(6);
      double A = ct * ct / (2.0 * sx * sx) + st * st / (2.0 * sy * sy);
      double B = sin(2.0 * th) / (2.0 * sx * sx) - sin(2.0 * th) / (2.0 * sy * sy);
      double Cc = st * st / (2.0 * sx * sx) + ct * ct / (2.0 * sy * sy);
      double dx = R(0) - R(3), dy = R(1) - R(4);
      *out = C(R(2) * exp(-(A * dx * dx + B * dx * dy + Cc * dy * dy)), 0.0);
      return 0;
    }
    case 3: {
      double phi = R(1) * sin(R(2) * R(0));
      *out = C(cos(phi), sin(phi));
      return 0;
    }
    case 4:
      *out = C(0.5 * (1.0 + (sin(R(1) * R(0) + R(2)) >= 0.0 ? 1.0 : -1.0)), 0.0);
      return 0;
    case 5: {
      double th = 0.5 * R(1) * sin(R(2) * R(0));
      double v = 0.5 * (cos(th) + sin(th));
      *out = C(v, v);
      return 0;
    }
    case 6:
      *out = C(rem_euclid((R(1) * R(0) + R(2)) / (2.0 * M_PI), 1.0), 0.0);
      return 0;
    case 7: {
      double period = R(1) + R(2), duty = R(1) / period;
      double frac = rem_euclid(2.0 * M_PI * R(0) / period + R(3), 2.0 * M_PI) / (2.0 * M_PI);
      *out = C(frac < duty ? 1.0 : 0.0, 0.0);
      return 0;
    }
    case 9:
      *out = C(rabi_from_intensity(R(0), R(1), argc > 2 ? R(2) : ED_XTB), 0.0);
      return 0;
    case 11: {
      double inten = gaussian_2d(R(0), R(1), R(2), R(3), R(4), R(5), R(6));
      *out = C(rabi_from_intensity(inten, R(7), argc > 8 ? R(8) : ED_XTB), 0.0);
      return 0;
    }
    case 12: {
      double frac = rem_euclid(R(0) * R(2) + R(3) / (2.0 * M_PI), 1.0);
      if (frac <= 0.0) frac += 1.0;
      *out = C(frac < R(1) ? 1.0 : 0.0, 0.0);
      return 0;
    }
    case 13: {
      long long k = (long long)floor((R(0) - R(1)) / R(2));
      *out = C(k % 2 == 0 ? 1.0 : -1.0, 0.0);
      return 0;
    }
    case 15: {
      double dx = R(0) - R(1);
      *out = C(exp(-dx * dx / (2.0 * R(2) * R(2))), 0.0);
      return 0;
    }
    default:
      return fail("helper function id %d is not supported by the scalar evaluator", id);
  }
#undef R
}

static int eval_expr(const instr_t* ins, const expr_t* e, const cpx* slots, double t, cpx* out) {
  cpx st[STACK_MAX];
  int sp = 0;
  for (int32_t p = e->start; p < e->end; ++p) {
    const instr_t* q = &ins[p];
    switch (q->op) {
      case OP_CONST: st[sp++] = C(q->re, q->im); break;
      case OP_SLOT: st[sp++] = slots[q->index]; break;
      case OP_TIME: st[sp++] = C(t, 0.0); break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_POW: {
        if (sp < 2) return fail("stack underflow on binary op");
        cpx r = st[--sp], l = st[--sp], v;
        if (q->op == OP_ADD) v = cadd(l, r);
        else if (q->op == OP_SUB) v = csub(l, r);
        else if (q->op == OP_MUL) v = cmul(l, r);
        else if (q->op == OP_DIV) v = cdiv(l, r);
        else v = cpow_nc(l, r);
        st[sp++] = v;
        break;
      }
      case OP_NEG: if (sp < 1) return fail("stack underflow on NEG"); st[sp - 1] = cneg(st[sp - 1]); break;
      case OP_CONJ: if (sp < 1) return fail("stack underflow on CONJ"); st[sp - 1] = cconj(st[sp - 1]); break;
      case OP_BUILTIN: {
        if (sp < q->argc || q->argc < 1) return fail("stack underflow on BUILTIN_FUNC");
        int start = sp - q->argc;
        cpx v = st[start], r;
        switch (q->function) {
          case 1: r = csin_nc(v); break;
          case 2: r = ccos_nc(v); break;
          case 3: r = ctan_nc(v); break;
          case 4: r = cexp_nc(v); break;
          case 5: r = C(hypot(v.re, v.im), 0.0); break;
          case 6: r = C(v.re, 0.0); break;
          case 7: r = C(v.im, 0.0); break;
          default: return fail("unknown builtin function id %d", q->function);
        }
        sp = start;
        st[sp++] = r;
        break;
      }
      case OP_HELPER: {
        if (sp < q->argc) return fail("stack underflow on HELPER_FUNC");
        int start = sp - q->argc;
        cpx r;
        int rc = apply_helper(q->function, &st[start], q->argc, &r);
        if (rc) return rc;
        sp = start;
        st[sp++] = r;
        break;
      }
      case OP_GT: case OP_GE: case OP_LT: case OP_LE: case OP_EQ: case OP_NE: {
        if (sp < 2) return fail("stack underflow on comparison");
        double r = st[--sp].re, l = st[--sp].re;
        int res = q->op == OP_GT ? l > r : q->op == OP_GE ? l >= r : q->op == OP_LT ? l < r
                : q->op == OP_LE ? l <= r : q->op == OP_EQ ? l == r : l != r;
        st[sp++] = C(res ? 1.0 : 0.0, 0.0);
        break;
      }
      default:
        return fail("unsupported instruction op %d", q->op);
    }
    if (sp >= STACK_MAX) return fail("expression stack overflow");
  }
  if (sp != 1) return fail("expression evaluation did not end with a single stack value");
  *out = st[0];
  return 0;
}

/* ---------------- plan --------------------------------------------------- */

typedef struct { int32_t re, im; double sign; int64_t src_start, src_end; } sterm_t;
typedef struct { int64_t term; int32_t re, im; double sign; } dterm_t;

struct lbo_plan {
  int n, dim, upper_len;
  int n_slots, n_base, n_compounds, n_coeff;
  cpx* base;
  int32_t* compound_slots;
  expr_t* exprs;
  instr_t* ins;
  int64_t n_terms;
  int64_t *out_ptrs, *in_idx, *coeff_idx;
  uint8_t *cconj, *iconj;
  cpx* factor;
  uint8_t* dyn_coeff;
  int time_dependent;
  int32_t *up_re, *up_im; /* im = -1 on the diagonal */
  int partitioned;
  /* partitioned (n <= 40) */
  int64_t *sr_ptr, *sc_ptr, *dr_ptr, *dc_ptr;
  sterm_t *sr, *sc;
  dterm_t *dr, *dc;
  int64_t* src; /* source term lists of static terms */
  int64_t n_sr, n_sc, n_dr, n_dc;
  /* split inputs (n > 40) */
  int64_t *gr_ptr, *gc_ptr;
  dterm_t *gr, *gc;
  /* effective-Hamiltonian plan (rust/src/effective_lindblad/): dim = 2 n^2 */
  int eff;
  int n_grid;
  double* grid;
  int64_t* e_rows[3];
  int64_t* e_cols[3];
  double* e_coef[3];
  int64_t e_nnz[3];
  int field_slot, rabi_slot, det_slot;
};

#define PARTITIONED_MAX_STATES 40

typedef struct { int64_t key0, key1, key2; int64_t term; } sort_item;

static int cmp_items(const void* a, const void* b) {
  const sort_item* x = a;
  const sort_item* y = b;
  if (x->key0 != y->key0) return x->key0 < y->key0 ? -1 : 1;
  if (x->key1 != y->key1) return x->key1 < y->key1 ? -1 : 1;
  if (x->key2 != y->key2) return x->key2 < y->key2 ? -1 : 1;
  if (x->term != y->term) return x->term < y->term ? -1 : 1; /* stable */
  return 0;
}

void lbo_plan_destroy(lbo_plan* p) {
  if (!p) return;
  void* ptrs[] = {p->base, p->compound_slots, p->exprs, p->ins, p->out_ptrs, p->in_idx,
                  p->coeff_idx, p->cconj, p->iconj, p->factor, p->dyn_coeff, p->up_re, p->up_im,
                  p->sr_ptr, p->sc_ptr, p->dr_ptr, p->dc_ptr, p->sr, p->sc, p->dr, p->dc, p->src,
                  p->gr_ptr, p->gc_ptr, p->gr, p->gc, p->grid,
                  p->e_rows[0], p->e_rows[1], p->e_rows[2], p->e_cols[0], p->e_cols[1], p->e_cols[2],
                  p->e_coef[0], p->e_coef[1], p->e_coef[2]};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(p);
}

int64_t lbo_plan_terms(const lbo_plan* p) { return p->n_terms; }

static int expr_reads_dynamic(const lbo_plan* p, const expr_t* e, const uint8_t* dyn_slot) {
  for (int32_t k = e->start; k < e->end; ++k) {
    if (p->ins[k].op == OP_TIME) return 1;
    if (p->ins[k].op == OP_SLOT && p->ins[k].index >= 0 && p->ins[k].index < p->n_slots &&
        dyn_slot[p->ins[k].index])
      return 1;
  }
  return 0;
}

int lbo_plan_create(const lb_plan_desc* d, lbo_plan** out) {
  *out = NULL;
  if (!d || d->n_states <= 0) return fail("n_states must be positive");
  lbo_plan* p = calloc(1, sizeof(*p));
  int n = d->n_states;
  p->n = n;
  p->dim = n * n;
  p->upper_len = n * (n + 1) / 2;
  p->n_slots = d->n_slots;
  p->n_base = d->n_base;
  p->n_compounds = d->n_compounds;
  p->n_coeff = d->n_coefficients;
  if (p->n_base > p->n_slots) { lbo_plan_destroy(p); return fail("n_base exceeds n_slots"); }
  p->base = calloc(p->n_base + 1, sizeof(cpx));
  for (int i = 0; i < p->n_base; ++i) p->base[i] = C(d->base_re[i], d->base_im[i]);
  p->compound_slots = calloc(p->n_compounds + 1, sizeof(int32_t));
  for (int i = 0; i < p->n_compounds; ++i) {
    p->compound_slots[i] = d->compound_slots[i];
    if (d->compound_slots[i] < p->n_base || d->compound_slots[i] >= p->n_slots) {
      lbo_plan_destroy(p);
      return fail("compound slot out of range");
    }
  }
  int n_expr = p->n_compounds + p->n_coeff;
  int32_t n_ins = d->expr_ptrs[n_expr];
  p->exprs = calloc(n_expr + 1, sizeof(expr_t));
  p->ins = calloc(n_ins + 1, sizeof(instr_t));
  for (int32_t i = 0; i < n_ins; ++i) {
    p->ins[i].op = d->instr_op[i];
    p->ins[i].index = d->instr_index[i];
    p->ins[i].argc = d->instr_argc[i];
    p->ins[i].function = d->instr_function[i];
    p->ins[i].re = d->instr_re[i];
    p->ins[i].im = d->instr_im[i];
    if (p->ins[i].op == OP_SLOT && (p->ins[i].index < 0 || p->ins[i].index >= p->n_slots)) {
      lbo_plan_destroy(p);
      return fail("slot index out of range in bytecode");
    }
  }
  for (int e = 0; e < n_expr; ++e) {
    p->exprs[e].start = d->expr_ptrs[e];
    p->exprs[e].end = d->expr_ptrs[e + 1];
    p->exprs[e].uses_time = 0;
    for (int32_t k = p->exprs[e].start; k < p->exprs[e].end; ++k)
      if (p->ins[k].op == OP_TIME) p->exprs[e].uses_time = 1;
  }
  /* dynamic classification (plan.rs:1235-1291) */
  uint8_t* dyn_slot = calloc(p->n_slots + 1, 1);
  p->time_dependent = 0;
  for (int c = 0; c < p->n_compounds; ++c) {
    dyn_slot[p->compound_slots[c]] = (uint8_t)expr_reads_dynamic(p, &p->exprs[c], dyn_slot);
    if (p->exprs[c].uses_time) p->time_dependent = 1;
  }
  p->dyn_coeff = calloc(p->n_coeff + 1, 1);
  for (int k = 0; k < p->n_coeff; ++k) {
    p->dyn_coeff[k] = (uint8_t)expr_reads_dynamic(p, &p->exprs[p->n_compounds + k], dyn_slot);
    if (p->exprs[p->n_compounds + k].uses_time) p->time_dependent = 1;
  }
  free(dyn_slot);

  /* expanded sparse program (plan.rs:1069-1153 checks) */
  int64_t T = d->n_terms;
  p->n_terms = T;
  if (d->output_ptrs[0] != 0 || d->output_ptrs[p->upper_len] != T) {
    lbo_plan_destroy(p);
    return fail("expanded RHS output_ptrs must start at 0 and end at term count");
  }
  p->out_ptrs = malloc(sizeof(int64_t) * (p->upper_len + 1));
  memcpy(p->out_ptrs, d->output_ptrs, sizeof(int64_t) * (p->upper_len + 1));
  for (int o = 0; o < p->upper_len; ++o)
    if (p->out_ptrs[o] > p->out_ptrs[o + 1]) {
      lbo_plan_destroy(p);
      return fail("expanded RHS output_ptrs must be monotonic");
    }
  p->in_idx = malloc(sizeof(int64_t) * (T + 1));
  p->coeff_idx = malloc(sizeof(int64_t) * (T + 1));
  p->cconj = malloc(T + 1);
  p->iconj = malloc(T + 1);
  p->factor = malloc(sizeof(cpx) * (T + 1));
  for (int64_t k = 0; k < T; ++k) {
    p->in_idx[k] = d->input_indices[k];
    p->coeff_idx[k] = d->coeff_indices[k];
    p->cconj[k] = d->coeff_conj[k];
    p->iconj[k] = d->input_conj[k];
    p->factor[k] = C(d->factors_re[k], d->factors_im[k]);
    if (p->in_idx[k] < 0 || p->in_idx[k] >= p->upper_len) {
      lbo_plan_destroy(p);
      return fail("expanded RHS input index out of bounds");
    }
    if (p->coeff_idx[k] >= p->n_coeff) {
      lbo_plan_destroy(p);
      return fail("expanded RHS coefficient index out of bounds");
    }
  }
  /* upper -> packed (rhs.rs:151-176) */
  p->up_re = malloc(sizeof(int32_t) * p->upper_len);
  p->up_im = malloc(sizeof(int32_t) * p->upper_len);
  {
    int cursor = n, u = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++u) {
        if (i == j) { p->up_re[u] = i; p->up_im[u] = -1; }
        else { p->up_re[u] = cursor; p->up_im[u] = cursor + 1; cursor += 2; }
      }
  }
  p->partitioned = n <= PARTITIONED_MAX_STATES;
  int U = p->upper_len;
  sort_item* items = malloc(sizeof(sort_item) * (T + 1));
  if (p->partitioned) {
    /* rhs.rs:285-384: per output, static terms grouped by packed input
       (BTreeMap order), dynamic terms sorted by packed input. */
    p->sr_ptr = calloc(U + 1, sizeof(int64_t));
    p->sc_ptr = calloc(U + 1, sizeof(int64_t));
    p->dr_ptr = calloc(U + 1, sizeof(int64_t));
    p->dc_ptr = calloc(U + 1, sizeof(int64_t));
    p->sr = malloc(sizeof(sterm_t) * (T + 1));
    p->sc = malloc(sizeof(sterm_t) * (T + 1));
    p->dr = malloc(sizeof(dterm_t) * (T + 1));
    p->dc = malloc(sizeof(dterm_t) * (T + 1));
    p->src = malloc(sizeof(int64_t) * (T + 1));
    int64_t nsrc = 0;
    for (int o = 0; o < U; ++o) {
      int64_t a = p->out_ptrs[o], b = p->out_ptrs[o + 1];
      /* static real */
      int64_t m = 0;
      for (int64_t k = a; k < b; ++k) {
        int dyn = p->coeff_idx[k] >= 0 && p->dyn_coeff[p->coeff_idx[k]];
        int32_t re = p->up_re[p->in_idx[k]], im = p->up_im[p->in_idx[k]];
        if (!dyn && im < 0) { items[m].key0 = re; items[m].key1 = 0; items[m].key2 = 0; items[m].term = k; ++m; }
      }
      qsort(items, m, sizeof(sort_item), cmp_items);
      for (int64_t q = 0; q < m;) {
        int64_t r = q;
        sterm_t s = {(int32_t)items[q].key0, (int32_t)items[q].key0, 0.0, nsrc, nsrc};
        while (r < m && items[r].key0 == items[q].key0) p->src[nsrc++] = items[r++].term;
        s.src_end = nsrc;
        p->sr[p->n_sr++] = s;
        q = r;
      }
      /* static complex: key (re, im, sign) */
      m = 0;
      for (int64_t k = a; k < b; ++k) {
        int dyn = p->coeff_idx[k] >= 0 && p->dyn_coeff[p->coeff_idx[k]];
        int32_t re = p->up_re[p->in_idx[k]], im = p->up_im[p->in_idx[k]];
        if (!dyn && im >= 0) {
          items[m].key0 = re; items[m].key1 = im; items[m].key2 = p->iconj[k] ? -1 : 1;
          items[m].term = k; ++m;
        }
      }
      qsort(items, m, sizeof(sort_item), cmp_items);
      for (int64_t q = 0; q < m;) {
        int64_t r = q;
        sterm_t s = {(int32_t)items[q].key0, (int32_t)items[q].key1, (double)items[q].key2, nsrc, nsrc};
        while (r < m && items[r].key0 == items[q].key0 && items[r].key1 == items[q].key1 &&
               items[r].key2 == items[q].key2)
          p->src[nsrc++] = items[r++].term;
        s.src_end = nsrc;
        p->sc[p->n_sc++] = s;
        q = r;
      }
      /* dynamic real: sorted by re; dynamic complex: sorted by (re, im) */
      m = 0;
      for (int64_t k = a; k < b; ++k) {
        int dyn = p->coeff_idx[k] >= 0 && p->dyn_coeff[p->coeff_idx[k]];
        int32_t re = p->up_re[p->in_idx[k]], im = p->up_im[p->in_idx[k]];
        if (dyn && im < 0) { items[m].key0 = re; items[m].key1 = 0; items[m].key2 = 0; items[m].term = k; ++m; }
      }
      qsort(items, m, sizeof(sort_item), cmp_items);
      for (int64_t q = 0; q < m; ++q) {
        dterm_t t = {items[q].term, (int32_t)items[q].key0, (int32_t)items[q].key0, 0.0};
        p->dr[p->n_dr++] = t;
      }
      m = 0;
      for (int64_t k = a; k < b; ++k) {
        int dyn = p->coeff_idx[k] >= 0 && p->dyn_coeff[p->coeff_idx[k]];
        int32_t re = p->up_re[p->in_idx[k]], im = p->up_im[p->in_idx[k]];
        if (dyn && im >= 0) { items[m].key0 = re; items[m].key1 = im; items[m].key2 = 0; items[m].term = k; ++m; }
      }
      qsort(items, m, sizeof(sort_item), cmp_items);
      for (int64_t q = 0; q < m; ++q) {
        int64_t k = items[q].term;
        dterm_t t = {k, (int32_t)items[q].key0, (int32_t)items[q].key1, p->iconj[k] ? -1.0 : 1.0};
        p->dc[p->n_dc++] = t;
      }
      p->sr_ptr[o + 1] = p->n_sr;
      p->sc_ptr[o + 1] = p->n_sc;
      p->dr_ptr[o + 1] = p->n_dr;
      p->dc_ptr[o + 1] = p->n_dc;
    }
  } else {
    /* rhs.rs:207-250: per output, real-input terms then complex-input terms in term order */
    p->gr_ptr = calloc(U + 1, sizeof(int64_t));
    p->gc_ptr = calloc(U + 1, sizeof(int64_t));
    p->gr = malloc(sizeof(dterm_t) * (T + 1));
    p->gc = malloc(sizeof(dterm_t) * (T + 1));
    int64_t nr = 0, nc = 0;
    for (int o = 0; o < U; ++o) {
      for (int64_t k = p->out_ptrs[o]; k < p->out_ptrs[o + 1]; ++k) {
        int32_t re = p->up_re[p->in_idx[k]], im = p->up_im[p->in_idx[k]];
        if (im < 0) { dterm_t t = {k, re, re, 0.0}; p->gr[nr++] = t; }
        else { dterm_t t = {k, re, im, p->iconj[k] ? -1.0 : 1.0}; p->gc[nc++] = t; }
      }
      p->gr_ptr[o + 1] = nr;
      p->gc_ptr[o + 1] = nc;
    }
  }
  free(items);
  *out = p;
  return LB_OK;
}

/* ---------------- workspace + RHS --------------------------------------- */

enum { CACHE_NONE = 0, CACHE_PARTITIONED = 1, CACHE_SPLIT = 2 };

typedef struct {
  const lbo_plan* p;
  cpx* slots;
  cpx* coeff;
  cpx* sr_val; /* folded static values, partitioned */
  cpx* sc_val;
  double *tv_re, *tv_im; /* per-term values */
  int n_override;
  const int32_t* ov_slots;
  const cpx* ov_values;
  int cache;
  int64_t eff_hint; /* the effective plan's cached grid interval (rhs.rs:10) */
} workspace;

static void ws_init(workspace* w, const lbo_plan* p) {
  memset(w, 0, sizeof(*w));
  w->p = p;
  w->slots = calloc(p->n_slots + 1, sizeof(cpx));
  w->coeff = calloc(p->n_coeff + 1, sizeof(cpx));
  w->sr_val = calloc(p->n_sr + 1, sizeof(cpx));
  w->sc_val = calloc(p->n_sc + 1, sizeof(cpx));
  w->tv_re = calloc(p->n_terms + 1, sizeof(double));
  w->tv_im = calloc(p->n_terms + 1, sizeof(double));
}

static void ws_free(workspace* w) {
  free(w->slots); free(w->coeff); free(w->sr_val); free(w->sc_val); free(w->tv_re); free(w->tv_im);
}

/* set_scalar_parameter_overrides (rhs.rs:1176-1197): invalidates the cache. */
static void ws_set_overrides(workspace* w, int n, const int32_t* slots, const cpx* values) {
  w->n_override = n;
  w->ov_slots = slots;
  w->ov_values = values;
  w->cache = CACHE_NONE;
  w->eff_hint = 0; /* a new trajectory starts a new interval hint */
}

/* evaluate_parameter_graph (rhs.rs:1176-1212): base values, overrides, compounds. */
static int eval_graph(workspace* w, double t) {
  const lbo_plan* p = w->p;
  for (int i = 0; i < p->n_slots; ++i) w->slots[i] = i < p->n_base ? p->base[i] : C(0.0, 0.0);
  for (int i = 0; i < w->n_override; ++i) {
    if (w->ov_slots[i] < 0 || w->ov_slots[i] >= p->n_base)
      return fail("parameter override slot %d is not a base parameter slot", w->ov_slots[i]);
    w->slots[w->ov_slots[i]] = w->ov_values[i];
  }
  for (int c = 0; c < p->n_compounds; ++c) {
    cpx v;
    int rc = eval_expr(p->ins, &p->exprs[c], w->slots, t, &v);
    if (rc) return rc;
    w->slots[p->compound_slots[c]] = v;
  }
  return 0;
}

static int eval_graph_and_coefficients(workspace* w, double t) {
  const lbo_plan* p = w->p;
  int rc0 = eval_graph(w, t);
  if (rc0) return rc0;
  for (int k = 0; k < p->n_coeff; ++k) {
    int rc = eval_expr(p->ins, &p->exprs[p->n_compounds + k], w->slots, t, &w->coeff[k]);
    if (rc) return rc;
  }
  return 0;
}

static inline cpx term_value(const lbo_plan* p, const workspace* w, int64_t k) {
  cpx c = C(1.0, 0.0);
  if (p->coeff_idx[k] >= 0) {
    c = w->coeff[p->coeff_idx[k]];
    if (p->cconj[k]) c = cconj(c);
  }
  return cmul(p->factor[k], c);
}

static inline void write_out(const lbo_plan* p, int o, double re, double im, double* out) {
  out[p->up_re[o]] = re;
  if (p->up_im[o] >= 0) out[p->up_im[o]] = im;
}

static int eff_rhs(workspace* w, const double* y, double t, double* out);

static int rhs_eval(workspace* w, const double* y, double t, double* out) {
  const lbo_plan* p = w->p;
  if (p->eff) return eff_rhs(w, y, t, out);
  if (p->partitioned) {
    int init_static = w->cache != CACHE_PARTITIONED;
    if (init_static || p->time_dependent) {
      int rc = eval_graph_and_coefficients(w, t);
      if (rc) return rc;
      if (init_static) {
        for (int64_t s = 0; s < p->n_sr; ++s) {
          cpx acc = C(0.0, 0.0);
          for (int64_t q = p->sr[s].src_start; q < p->sr[s].src_end; ++q)
            acc = cadd(acc, term_value(p, w, p->src[q]));
          w->sr_val[s] = acc;
        }
        for (int64_t s = 0; s < p->n_sc; ++s) {
          cpx acc = C(0.0, 0.0);
          for (int64_t q = p->sc[s].src_start; q < p->sc[s].src_end; ++q)
            acc = cadd(acc, term_value(p, w, p->src[q]));
          w->sc_val[s] = acc;
        }
      }
      for (int64_t s = 0; s < p->n_dr; ++s) {
        cpx v = term_value(p, w, p->dr[s].term);
        w->tv_re[p->dr[s].term] = v.re;
        w->tv_im[p->dr[s].term] = v.im;
      }
      for (int64_t s = 0; s < p->n_dc; ++s) {
        cpx v = term_value(p, w, p->dc[s].term);
        w->tv_re[p->dc[s].term] = v.re;
        w->tv_im[p->dc[s].term] = v.im;
      }
      w->cache = CACHE_PARTITIONED;
    }
    for (int o = 0; o < p->upper_len; ++o) {
      double ar = 0.0, ai = 0.0;
      for (int64_t s = p->sr_ptr[o]; s < p->sr_ptr[o + 1]; ++s) {
        double vr = y[p->sr[s].re];
        ar += w->sr_val[s].re * vr;
        ai += w->sr_val[s].im * vr;
      }
      for (int64_t s = p->sc_ptr[o]; s < p->sc_ptr[o + 1]; ++s) {
        double vr = y[p->sc[s].re], vi = p->sc[s].sign * y[p->sc[s].im];
        ar += w->sc_val[s].re * vr - w->sc_val[s].im * vi;
        ai += w->sc_val[s].re * vi + w->sc_val[s].im * vr;
      }
      for (int64_t s = p->dr_ptr[o]; s < p->dr_ptr[o + 1]; ++s) {
        double vr = y[p->dr[s].re];
        ar += w->tv_re[p->dr[s].term] * vr;
        ai += w->tv_im[p->dr[s].term] * vr;
      }
      for (int64_t s = p->dc_ptr[o]; s < p->dc_ptr[o + 1]; ++s) {
        double cr = w->tv_re[p->dc[s].term], ci = w->tv_im[p->dc[s].term];
        double vr = y[p->dc[s].re], vi = p->dc[s].sign * y[p->dc[s].im];
        ar += cr * vr - ci * vi;
        ai += cr * vi + ci * vr;
      }
      write_out(p, o, ar, ai, out);
    }
    return 0;
  }
  int can_skip = !p->time_dependent && w->cache == CACHE_SPLIT;
  if (!can_skip) {
    int rc = eval_graph_and_coefficients(w, t);
    if (rc) return rc;
    for (int64_t k = 0; k < p->n_terms; ++k) {
      cpx v = term_value(p, w, k);
      w->tv_re[k] = v.re;
      w->tv_im[k] = v.im;
    }
    w->cache = CACHE_SPLIT;
  }
  for (int o = 0; o < p->upper_len; ++o) {
    double ar = 0.0, ai = 0.0;
    for (int64_t s = p->gr_ptr[o]; s < p->gr_ptr[o + 1]; ++s) {
      double vr = y[p->gr[s].re];
      ar += w->tv_re[p->gr[s].term] * vr;
      ai += w->tv_im[p->gr[s].term] * vr;
    }
    for (int64_t s = p->gc_ptr[o]; s < p->gc_ptr[o + 1]; ++s) {
      double cr = w->tv_re[p->gc[s].term], ci = w->tv_im[p->gc[s].term];
      double vr = y[p->gc[s].re], vi = p->gc[s].sign * y[p->gc[s].im];
      ar += cr * vr - ci * vi;
      ai += cr * vi + ci * vr;
    }
    write_out(p, o, ar, ai, out);
  }
  return 0;
}

/* ---------------- effective-Hamiltonian plans ------------------------------ */

/* Rust's {:.6e}: six fraction digits, exponent without sign padding. */
static void rust_exp6(double x, char* buf, size_t size) {
  char tmp[64];
  snprintf(tmp, sizeof(tmp), "%.6e", x);
  char* e = strchr(tmp, 'e');
  if (!e) { snprintf(buf, size, "%s", tmp); return; }
  *e = 0;
  int ex = atoi(e + 1);
  snprintf(buf, size, "%se%d", tmp, ex);
}

/* find_interval_cached (rust/src/effective_lindblad/rhs.rs:47-84). */
static int eff_find_interval(const lbo_plan* p, double x, int64_t* hint, int64_t* idx_out) {
  const int64_t n = p->n_grid;
  const double* g = p->grid;
  if (n < 2) { *idx_out = 0; return 0; }
  const double tol = 1e-10 * fmax(fabs(g[n - 1] - g[0]), 1.0);
  char a[64], b[64];
  if (x < g[0] - tol) {
    rust_exp6(x, a, sizeof(a)); rust_exp6(g[0], b, sizeof(b));
    fail("field coordinate %s is below the operator grid minimum %s", a, b);
    return LB_TRAJ_FIELD_BELOW;
  }
  if (x > g[n - 1] + tol) {
    rust_exp6(x, a, sizeof(a)); rust_exp6(g[n - 1], b, sizeof(b));
    fail("field coordinate %s is above the operator grid maximum %s", a, b);
    return LB_TRAJ_FIELD_ABOVE;
  }
  const double xc = x < g[0] ? g[0] : (x > g[n - 1] ? g[n - 1] : x);
  if (xc <= g[0]) { *hint = 0; *idx_out = 0; return 0; }
  if (xc >= g[n - 1]) { *hint = n - 2; *idx_out = n - 2; return 0; }
  const int64_t h = *hint;
  if (h < n - 1 && g[h] <= xc && xc <= g[h + 1]) { *idx_out = h; return 0; }
  int64_t count = 0; /* partition_point(|v| v <= xc) */
  while (count < n && g[count] <= xc) ++count;
  int64_t idx = count > 0 ? count - 1 : 0;
  if (idx > n - 2) idx = n - 2;
  *hint = idx;
  *idx_out = idx;
  return 0;
}

/* rhs_effective_lindblad (rhs.rs:86-138): dy = 0; then each operator's
   interpolated CSR row sums scaled by 1, Omega/2, delta (plan.rs:18-40). */
static int eff_rhs(workspace* w, const double* y, double t, double* out) {
  const lbo_plan* p = w->p;
  int rc = eval_graph(w, t);
  if (rc) return rc;
  const double x = w->slots[p->field_slot].re;
  const double rabi = w->slots[p->rabi_slot].re;
  const double det = w->slots[p->det_slot].re;
  int64_t idx = 0;
  if ((rc = eff_find_interval(p, x, &w->eff_hint, &idx))) return rc;
  const double dx = x - p->grid[idx];
  const double scale[3] = {1.0, 0.5 * rabi, det};
  for (int i = 0; i < p->dim; ++i) out[i] = 0.0;
  for (int o = 0; o < 3; ++o) {
    const int64_t interval = idx < p->n_grid - 2 ? idx : p->n_grid - 2;
    const double* cb = p->e_coef[o] + interval * p->e_nnz[o] * 4;
    for (int i = 0; i < p->dim; ++i) {
      double sum = 0.0;
      for (int64_t q = p->e_rows[o][i]; q < p->e_rows[o][i + 1]; ++q) {
        const double* c = cb + q * 4;
        const double val = c[0] + dx * (c[1] + dx * (c[2] + dx * c[3]));
        sum += val * y[p->e_cols[o][q]];
      }
      out[i] += scale[o] * sum;
    }
  }
  return 0;
}

int lbo_eff_plan_create(const lb_eff_plan_desc* d, lbo_plan** out) {
  *out = NULL;
  if (!d || d->n_states <= 0) return fail("n_states must be positive");
  if (d->real_dim != 2 * d->n_states * d->n_states) return fail("real_dim must be 2 * n_states^2");
  if (d->n_grid < 2) return fail("the operator grid needs at least two points");
  lbo_plan* p = calloc(1, sizeof(*p));
  p->eff = 1;
  p->n = d->n_states;
  p->dim = d->real_dim;
  p->n_slots = d->n_slots;
  p->n_base = d->n_base;
  p->n_compounds = d->n_compounds;
  p->n_grid = d->n_grid;
  p->field_slot = d->field_slot;
  p->rabi_slot = d->rabi_slot;
  p->det_slot = d->detuning_slot;
  p->grid = malloc(sizeof(double) * d->n_grid);
  memcpy(p->grid, d->field_grid, sizeof(double) * d->n_grid);
  for (int o = 0; o < 3; ++o) {
    const int64_t nnz = d->nnz[o];
    p->e_nnz[o] = nnz;
    p->e_rows[o] = malloc(sizeof(int64_t) * (p->dim + 1));
    memcpy(p->e_rows[o], d->row_ptrs[o], sizeof(int64_t) * (p->dim + 1));
    p->e_cols[o] = malloc(sizeof(int64_t) * (nnz + 1));
    if (nnz) memcpy(p->e_cols[o], d->col_indices[o], sizeof(int64_t) * nnz);
    const size_t nc = (size_t)(d->n_grid - 1) * nnz * 4;
    p->e_coef[o] = malloc(sizeof(double) * (nc + 1));
    if (nc) memcpy(p->e_coef[o], d->coeffs[o], sizeof(double) * nc);
  }
  p->base = calloc(p->n_base + 1, sizeof(cpx));
  for (int i = 0; i < p->n_base; ++i) p->base[i] = C(d->base_re[i], d->base_im[i]);
  p->compound_slots = calloc(p->n_compounds + 1, sizeof(int32_t));
  p->exprs = calloc(p->n_compounds + 1, sizeof(expr_t));
  const int32_t n_ins = p->n_compounds ? d->expr_ptrs[p->n_compounds] : 0;
  p->ins = calloc(n_ins + 1, sizeof(instr_t));
  for (int32_t q = 0; q < n_ins; ++q)
    p->ins[q] = (instr_t){d->instr_op[q], d->instr_index[q], d->instr_argc[q], d->instr_function[q], d->instr_re[q],
                          d->instr_im[q]};
  for (int c = 0; c < p->n_compounds; ++c) {
    p->compound_slots[c] = d->compound_slots[c];
    p->exprs[c].start = d->expr_ptrs[c];
    p->exprs[c].end = d->expr_ptrs[c + 1];
  }
  *out = p;
  return LB_OK;
}

int lbo_rhs(lbo_plan* plan, const double* y, double t, int32_t n_override,
            const int32_t* override_slots, const double* override_values, double* out) {
  workspace w;
  ws_init(&w, plan);
  ws_set_overrides(&w, n_override, override_slots, (const cpx*)override_values);
  int rc = rhs_eval(&w, y, t, out);
  ws_free(&w);
  return rc;
}

/* ---------------- collectors (output.rs) -------------------------------- */

typedef struct {
  const lbo_solve_args* a;
  const lbo_plan* p;
  int64_t capacity, width, count;
  double* times;
  double* values;
  /* integral state */
  double integral, last_t, last_value;
  /* the weighted trapezoid over t0 (save_start) and every accepted step,
   * collected alongside another output (lbo_solve_args.integral_out) */
  double side_integral, side_last_t, side_last_value;
  double last_push_t; /* time of the latest push (any output kind), NaN before the first */
} collector;

static double weighted(const collector* c, const double* y) {
  double s = 0.0;
  for (int i = 0; i < c->a->n_weights; ++i) s += c->a->weight_values[i] * y[c->a->weight_indices[i]];
  return s;
}

static int64_t packed_upper_real_index(int n, int i, int j) {
  int64_t off = 0;
  for (int r = 0; r < i; ++r) off += n - r - 1;
  return n + 2 * (off + (j - i - 1));
}

static void coll_push(collector* c, double t, const double* y) {
  const lbo_solve_args* a = c->a;
  int n = c->p->n;
  c->last_push_t = t;
  if (a->output_kind == LBO_OUT_INTEGRAL) {
    double v = weighted(c, y);
    if (isfinite(c->last_t)) c->integral += 0.5 * (c->last_value + v) * (t - c->last_t);
    c->last_t = t;
    c->last_value = v;
    if (a->store_trace) {
      if (c->count < c->capacity) {
        c->times[c->count] = t;
        c->values[c->count] = c->integral;
      }
      c->count++;
    } else {
      c->times[0] = t;
      c->values[0] = c->integral;
      c->count = 1;
    }
    return;
  }
  if (c->count >= c->capacity) { c->count++; return; }
  double* row = c->values + c->count * c->width;
  c->times[c->count] = t;
  if (c->p->eff && a->output_kind == LBO_OUT_POPULATIONS) {
    /* split_real_diagonal_indices (solver.rs:56-58) */
    for (int i = 0; i < n; ++i) row[i] = y[i * n + i];
    c->count++;
    return;
  }
  if (c->p->eff && a->output_kind == LBO_OUT_SELECTED) {
    /* split_real_selected_extractions (solver.rs:60-76) */
    for (int s = 0; s < a->n_selected; ++s) {
      int i = a->selected_rows[s], j = a->selected_cols[s];
      row[2 * s] = y[i * n + j];
      row[2 * s + 1] = i == j ? 0.0 : y[n * n + i * n + j];
    }
    c->count++;
    return;
  }
  switch (a->output_kind) {
    case LBO_OUT_POPULATIONS:
      for (int i = 0; i < n; ++i) row[i] = y[i];
      break;
    case LBO_OUT_SELECTED:
      for (int s = 0; s < a->n_selected; ++s) {
        int i = a->selected_rows[s], j = a->selected_cols[s];
        if (i == j) { row[2 * s] = y[i]; row[2 * s + 1] = 0.0; }
        else if (i < j) { int64_t r = packed_upper_real_index(n, i, j); row[2 * s] = y[r]; row[2 * s + 1] = y[r + 1]; }
        else { int64_t r = packed_upper_real_index(n, j, i); row[2 * s] = y[r]; row[2 * s + 1] = -y[r + 1]; }
      }
      break;
    case LBO_OUT_FULL:
      memcpy(row, y, sizeof(double) * c->p->dim);
      break;
    case LBO_OUT_RATE:
      row[0] = weighted(c, y);
      break;
  }
  c->count++;
}

/* The integral solve the reference would run beside this one (saveat None):
 * pushes at t0 when saving the start and at every accepted step. */
static void side_push(collector* c, double t, const double* y) {
  if (!c->a->integral_out) return;
  double v = weighted(c, y);
  if (isfinite(c->side_last_t)) c->side_integral += 0.5 * (c->side_last_value + v) * (t - c->side_last_t);
  c->side_last_t = t;
  c->side_last_value = v;
}

int64_t lbo_output_width(const lbo_plan* p, const lbo_solve_args* a) {
  switch (a->output_kind) {
    case LBO_OUT_POPULATIONS: return p->n;
    case LBO_OUT_SELECTED: return 2 * (int64_t)a->n_selected;
    case LBO_OUT_FULL: return p->dim;
    default: return 1;
  }
}

/* ---------------- integrators (common.rs, dopri5.rs, tsit5.rs) ----------- */

static inline double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }
static inline double time_tol(double t) { return 10.0 * DBL_EPSILON * fmax(fabs(t), 1.0); }
static inline int close_to_start(double t, double t0) { return fabs(t - t0) <= time_tol(t0); }

/* ---------------- terminal stop events ------------------------------------ */

#define EVENT_COMPLEX (-9)

/* event_value (rust/src/lindblad/ode_impl.rs:134-153): the listed packed
   entries (the populations) summed in list order minus the threshold, or the
   real part of the event expression over the parameter graph at t
   (rhs.rs:1214-1240, an error beyond |Im| 1e-12). */
static int event_value(workspace* w, const lbo_solve_args* a, const instr_t* ev_ins, double t, const double* y,
                       double* g) {
  if (a->event_kind == LB_EVENT_POPULATION) {
    double v = 0.0;
    for (int k = 0; k < a->n_event_indices; ++k) v += y[a->event_indices[k]];
    *g = v - a->event_threshold;
    return 0;
  }
  int rc = eval_graph(w, t);
  if (rc) return rc;
  expr_t e = {0, a->n_event_instr, 1};
  cpx v;
  if ((rc = eval_expr(ev_ins, &e, w->slots, t, &v))) return rc;
  if (fabs(v.im) > 1e-12) {
    fail("event expression evaluated to a complex value");
    return EVENT_COMPLEX;
  }
  *g = v.re;
  return 0;
}

/* crossed_zero (dopri5.rs:145-147, fixed.rs:12-14) */
static inline int crossed_zero(double g0, double g1) {
  return g0 == 0.0 || g1 == 0.0 || (g0 < 0.0 && g1 > 0.0) || (g0 > 0.0 && g1 < 0.0);
}

/* The event program as oracle instructions (NULL for population events). */
static instr_t* event_program(const lbo_solve_args* a) {
  if (a->event_kind != LB_EVENT_EXPRESSION) return NULL;
  instr_t* ins = calloc((size_t)a->n_event_instr + 1, sizeof(instr_t));
  for (int i = 0; i < a->n_event_instr; ++i)
    ins[i] = (instr_t){a->event_op[i], a->event_index[i], a->event_argc[i], a->event_function[i], a->event_re[i],
                       a->event_im[i]};
  return ins;
}

typedef struct {
  double alpha, beta, facc1, facc2, fac_old, h_max, safety, posneg;
  int reject;
} controller;

static void ctrl_init(controller* c, double t0, double t1) {
  c->beta = 0.04;
  c->alpha = 0.2 - c->beta * 0.75;
  c->facc1 = 1.0 / 0.2;
  c->facc2 = 1.0 / 10.0;
  c->fac_old = 1.0e-4;
  c->h_max = fabs(t1 - t0);
  c->reject = 0;
  c->safety = 0.9;
  c->posneg = sign_of(1.0, t1 - t0);
}

/* Rust f64::max/min ignore a NaN operand; fmax/fmin have the same rule. */
static int ctrl_accept(controller* c, double err, double h, double* h_new_out) {
  double fac11 = pow(err, c->alpha);
  double fac = fac11 * pow(c->fac_old, -c->beta);
  fac = fmax(c->facc2, fmin(c->facc1, fac / c->safety));
  double h_new = h / fac;
  if (err <= 1.0) {
    c->fac_old = fmax(err, 1.0e-4);
    if (fabs(h_new) > c->h_max) h_new = c->posneg * c->h_max;
    if (c->reject) h_new = c->posneg * fmin(fabs(h_new), fabs(h));
    c->reject = 0;
    *h_new_out = h_new;
    return 1;
  }
  h_new = h / fmin(c->facc1, fac11 / c->safety);
  c->reject = 1;
  *h_new_out = h_new;
  return 0;
}

typedef struct {
  double c[8];
  double a[8][8];
  double e[8];
  int is_tsit5;
} tableau;

static tableau dopri5_tab(void) {
  tableau t;
  memset(&t, 0, sizeof(t));
  t.c[2] = 1.0 / 5.0; t.c[3] = 3.0 / 10.0; t.c[4] = 4.0 / 5.0; t.c[5] = 8.0 / 9.0; t.c[6] = 1.0; t.c[7] = 1.0;
  t.a[2][1] = 1.0 / 5.0;
  t.a[3][1] = 3.0 / 40.0; t.a[3][2] = 9.0 / 40.0;
  t.a[4][1] = 44.0 / 45.0; t.a[4][2] = -56.0 / 15.0; t.a[4][3] = 32.0 / 9.0;
  t.a[5][1] = 19372.0 / 6561.0; t.a[5][2] = -25360.0 / 2187.0; t.a[5][3] = 64448.0 / 6561.0; t.a[5][4] = -212.0 / 729.0;
  t.a[6][1] = 9017.0 / 3168.0; t.a[6][2] = -355.0 / 33.0; t.a[6][3] = 46732.0 / 5247.0; t.a[6][4] = 49.0 / 176.0; t.a[6][5] = -5103.0 / 18656.0;
  t.a[7][1] = 35.0 / 384.0; t.a[7][2] = 0.0; t.a[7][3] = 500.0 / 1113.0; t.a[7][4] = 125.0 / 192.0; t.a[7][5] = -2187.0 / 6784.0; t.a[7][6] = 11.0 / 84.0;
  t.e[1] = 71.0 / 57600.0; t.e[2] = 0.0; t.e[3] = -71.0 / 16695.0; t.e[4] = 71.0 / 1920.0; t.e[5] = -17253.0 / 339200.0; t.e[6] = 22.0 / 525.0; t.e[7] = -1.0 / 40.0;
  t.is_tsit5 = 0;
  return t;
}

static tableau tsit5_tab(void) {
  tableau t;
  memset(&t, 0, sizeof(t));
  t.c[2] = 0.161; t.c[3] = 0.327; t.c[4] = 0.9; t.c[5] = 0.9800255409045097; t.c[6] = 1.0; t.c[7] = 1.0;
  t.a[2][1] = 0.161;
  t.a[3][1] = -0.008480655492356989; t.a[3][2] = 0.335480655492357;
  t.a[4][1] = 2.8971530571054935; t.a[4][2] = -6.359448489975075; t.a[4][3] = 4.3622954328695815;
  t.a[5][1] = 5.325864828439257; t.a[5][2] = -11.748883564062828; t.a[5][3] = 7.4955393428898365; t.a[5][4] = -0.09249506636175525;
  t.a[6][1] = 5.86145544294642; t.a[6][2] = -12.92096931784711; t.a[6][3] = 8.159367898576159; t.a[6][4] = -0.071584973281401; t.a[6][5] = -0.028269050394068383;
  t.a[7][1] = 0.09646076681806523; t.a[7][2] = 0.01; t.a[7][3] = 0.4798896504144996; t.a[7][4] = 1.379008574103742; t.a[7][5] = -3.290069515436081; t.a[7][6] = 2.324710524099774;
  t.e[1] = -0.0017800110522257771; t.e[2] = -0.0008164344596567469; t.e[3] = 0.007880878010261995; t.e[4] = -0.1447110071732629; t.e[5] = 0.5823571654525552; t.e[6] = -0.458082105929187; t.e[7] = 1.0 / 66.0;
  t.is_tsit5 = 1;
  return t;
}

/* Dense output coefficients (dopri5.rs:38-43, 104-143; tsit5.rs:39-60, 121-142). */
static const double DP_D[8] = {0, -12715105075.0 / 11282082432.0, 0, 87487479700.0 / 32700410799.0,
                               -10690763975.0 / 1880347072.0, 701980252875.0 / 199316789632.0,
                               -1453857185.0 / 822651844.0, 69997945.0 / 29380423.0};

static void dense_dopri5(double th, double h, int dim, const double* y, const double* yn,
                         double* const* k, double* out) {
  double th1 = 1.0 - th;
  for (int i = 0; i < dim; ++i) {
    double d = yn[i] - y[i];
    double b = k[1][i] * h - d;
    double r3 = -k[7][i] * h + d - b;
    double r4 = h * (DP_D[1] * k[1][i] + DP_D[3] * k[3][i] + DP_D[4] * k[4][i] + DP_D[5] * k[5][i] +
                     DP_D[6] * k[6][i] + DP_D[7] * k[7][i]);
    out[i] = y[i] + (d + (b + (r3 + r4 * th1) * th) * th1) * th;
  }
}

static void dense_tsit5(double th, double h, int dim, const double* y, double* const* k, double* out) {
  double t2 = th * th, t3 = t2 * th, t4 = t3 * th;
  double b1 = 1.0 * th + -2.763706197274826 * t2 + 2.9132554618219126 * t3 + -1.0530884977290216 * t4;
  double b2 = 0.1317 * t2 + -0.2234 * t3 + 0.1017 * t4;
  double b3 = 3.9302962368947516 * t2 + -5.941033872131505 * t3 + 2.490627285651253 * t4;
  double b4 = -12.411077166933676 * t2 + 30.33818863028232 * t3 + -16.548102889244902 * t4;
  double b5 = 37.50931341651104 * t2 + -88.1789048947664 * t3 + 47.37952196281928 * t4;
  double b6 = -27.896526289197286 * t2 + 65.09189467479366 * t3 + -34.87065786149661 * t4;
  double b7 = 1.5 * t2 + -4.0 * t3 + 2.5 * t4;
  for (int i = 0; i < dim; ++i)
    out[i] = y[i] + h * (b1 * k[1][i] + b2 * k[2][i] + b3 * k[3][i] + b4 * k[4][i] + b5 * k[5][i] +
                         b6 * k[6][i] + b7 * k[7][i]);
}

typedef struct {
  int64_t accepted, rejected, rhs_calls;
  int status;
  double fail_x;
  int event_fired;
  double event_t;
} ode_stats;

/* Stage input (dopri5.rs:45-86, tsit5.rs:62-102): the first stage is
   y + h*a21*k1 evaluated left to right, later ones y + h*(sum_j a_sj k_j). */
static void fill_stage(const tableau* tb, int s, int dim, const double* y, double h, double* const* k,
                       double* out) {
  if (s == 2) {
    for (int i = 0; i < dim; ++i) out[i] = y[i] + h * tb->a[2][1] * k[1][i];
    return;
  }
  for (int i = 0; i < dim; ++i) {
    double acc = tb->a[s][1] * k[1][i];
    for (int j = 2; j < s; ++j) acc = acc + tb->a[s][j] * k[j][i];
    out[i] = y[i] + h * acc;
  }
}

/* ---------------- fixed-step solvers (rust/src/ode/fixed.rs:133-249) ------ */
/* Steps of dt, shortened to land on t1 and on the next save-plan point, with
 * the stage sums written in the reference's operation order; no error
 * control.  Pushes after every step, or at the save-plan points (which the
 * steps land on). */
static int solve_fixed(const lbo_plan* p, workspace* w, const lbo_solve_args* a, const double* y0,
                       double t0, double t1, collector* out, ode_stats* st) {
  int dim = p->dim;
  memset(st, 0, sizeof(*st));
  if (a->maxiters <= 0) return fail("maxiters must be positive");
  if (t1 < t0) { st->status = LB_TRAJ_BAD_SPAN; st->fail_x = t0; return fail("only forward integration"); }
  if (!isfinite(a->dt) || a->dt <= 0.0) {
    st->status = LB_TRAJ_BAD_DT; st->fail_x = t0;
    return fail("fixed-step solvers require positive finite dt");
  }
  if (t1 == t0) {
    if (a->save_start) coll_push(out, t0, y0);
    return 0;
  }
  double* sp = NULL;
  int64_t nsp = 0;
  int has_sp = a->saveat != NULL || a->saveat_is_t1;
  if (a->saveat_is_t1) {
    /* output_when="final": the save plan {t1} of this trajectory (python_api.rs:589-617) */
    sp = malloc(sizeof(double));
    sp[nsp++] = t1;
  } else if (has_sp) {
    sp = malloc(sizeof(double) * (a->n_saveat + 1));
    if (a->save_start) sp[nsp++] = t0;
    for (int64_t q = 0; q < a->n_saveat; ++q) {
      double tt = a->saveat[q];
      if (a->save_start && close_to_start(tt, t0)) continue;
      if (!isfinite(tt)) { free(sp); return fail("saveat contains non-finite value"); }
      if (tt < t0 - time_tol(t0) || tt > t1 + time_tol(t1)) {
        free(sp);
        return fail("saveat value %g is outside [%g, %g]", tt, t0, t1);
      }
      sp[nsp++] = tt;
    }
  }
  double* buf = calloc((size_t)dim * 9, sizeof(double));
  instr_t* ev_ins = event_program(a);
  double *y = buf, *yn = buf + dim, *yt = buf + 2 * dim;
  double *k1 = buf + 3 * dim, *k2 = buf + 4 * dim, *k3 = buf + 5 * dim, *k4 = buf + 6 * dim,
         *k5 = buf + 7 * dim, *k6 = buf + 8 * dim;
  memcpy(y, y0, sizeof(double) * dim);
  int rc = 0;
  int64_t si = 0, steps = 0;
  double x = t0;
  if (has_sp) {
    while (si < nsp && close_to_start(sp[si], t0)) coll_push(out, sp[si++], y);
  } else if (a->save_start) {
    coll_push(out, x, y);
  }
  double ev_old = 0.0;
  if (a->event_kind && (rc = event_value(w, a, ev_ins, x, y, &ev_old))) {
    if (rc == EVENT_COMPLEX) { st->status = LB_TRAJ_EVENT_COMPLEX; st->fail_x = x; }
    goto done;
  }
  while (x < t1 - time_tol(t1)) {
    if (steps >= a->maxiters) {
      st->status = LB_TRAJ_MAXITERS; st->fail_x = x;
      rc = fail("Stopped at x=%.17g. Need more than %lld steps.", x, (long long)steps);
      goto done;
    }
    double target = fmin(x + a->dt, t1);
    if (has_sp && si < nsp && sp[si] > x + time_tol(x)) target = fmin(target, sp[si]);
    const double h = target - x;
    if (h <= 0.0 || 0.1 * h <= DBL_EPSILON * fmax(fabs(x), 1.0)) {
      st->status = LB_TRAJ_UNDERFLOW; st->fail_x = x;
      rc = fail("Stopped at x=%.17g. Step size underflow.", x);
      goto done;
    }
    int calls = 0;
    if (a->solver == LB_SOLVER_FIXED_RK2) {
      if ((rc = rhs_eval(w, y, x, k1))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + 0.5 * h * k1[i];
      if ((rc = rhs_eval(w, yt, x + 0.5 * h, k2))) goto done;
      for (int i = 0; i < dim; ++i) yn[i] = y[i] + h * k2[i];
      calls = 2;
    } else if (a->solver == LB_SOLVER_FIXED_RK4) {
      if ((rc = rhs_eval(w, y, x, k1))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + 0.5 * h * k1[i];
      if ((rc = rhs_eval(w, yt, x + 0.5 * h, k2))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + 0.5 * h * k2[i];
      if ((rc = rhs_eval(w, yt, x + 0.5 * h, k3))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + h * k3[i];
      if ((rc = rhs_eval(w, yt, x + h, k4))) goto done;
      for (int i = 0; i < dim; ++i) yn[i] = y[i] + h * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
      calls = 4;
    } else {
      if ((rc = rhs_eval(w, y, x, k1))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + h * (1.0 / 5.0) * k1[i];
      if ((rc = rhs_eval(w, yt, x + h * (1.0 / 5.0), k2))) goto done;
      for (int i = 0; i < dim; ++i) yt[i] = y[i] + h * ((3.0 / 40.0) * k1[i] + (9.0 / 40.0) * k2[i]);
      if ((rc = rhs_eval(w, yt, x + h * (3.0 / 10.0), k3))) goto done;
      for (int i = 0; i < dim; ++i)
        yt[i] = y[i] + h * ((44.0 / 45.0) * k1[i] - (56.0 / 15.0) * k2[i] + (32.0 / 9.0) * k3[i]);
      if ((rc = rhs_eval(w, yt, x + h * (4.0 / 5.0), k4))) goto done;
      for (int i = 0; i < dim; ++i)
        yt[i] = y[i] + h * ((19372.0 / 6561.0) * k1[i] - (25360.0 / 2187.0) * k2[i] +
                            (64448.0 / 6561.0) * k3[i] - (212.0 / 729.0) * k4[i]);
      if ((rc = rhs_eval(w, yt, x + h * (8.0 / 9.0), k5))) goto done;
      for (int i = 0; i < dim; ++i)
        yt[i] = y[i] + h * ((9017.0 / 3168.0) * k1[i] - (355.0 / 33.0) * k2[i] + (46732.0 / 5247.0) * k3[i] +
                            (49.0 / 176.0) * k4[i] - (5103.0 / 18656.0) * k5[i]);
      if ((rc = rhs_eval(w, yt, x + h, k6))) goto done;
      for (int i = 0; i < dim; ++i)
        yn[i] = y[i] + h * ((35.0 / 384.0) * k1[i] + (500.0 / 1113.0) * k3[i] + (125.0 / 192.0) * k4[i] -
                            (2187.0 / 6784.0) * k5[i] + (11.0 / 84.0) * k6[i]);
      calls = 6;
    }
    st->rhs_calls += calls;
    st->accepted += 1;
    steps += 1;
    const double xo = x;
    x += h;
    /* stop event (fixed.rs:213-238): the root is the step end, or its start
       when g was zero there, with the end state */
    int ev_hit = 0;
    if (a->event_kind) {
      double g1;
      if ((rc = event_value(w, a, ev_ins, x, yn, &g1))) {
        if (rc == EVENT_COMPLEX) { st->status = LB_TRAJ_EVENT_COMPLEX; st->fail_x = x; }
        goto done;
      }
      if (crossed_zero(ev_old, g1)) {
        ev_hit = 1;
        st->event_fired = 1;
        st->event_t = ev_old == 0.0 ? xo : x;
      }
      ev_old = g1;
    }
    if (has_sp) {
      const double stop = ev_hit ? st->event_t : x;
      while (si < nsp && sp[si] <= stop + time_tol(stop)) coll_push(out, sp[si++], yn);
    } else if (!ev_hit) {
      coll_push(out, x, yn);
    }
    if (ev_hit) {
      coll_push(out, st->event_t, yn);
      break;
    }
    double* tmp = y; y = yn; yn = tmp;
  }
  if (has_sp && si != nsp && !st->event_fired) {
    st->status = LB_TRAJ_SAVEAT; st->fail_x = (double)si;
    rc = fail("ended before all saveat: %lld of %lld", (long long)si, (long long)nsp);
  }
done:
  if (rc && st->status == LB_TRAJ_OK) st->status = 5;
  free(buf);
  free(sp);
  free(ev_ins);
  return rc;
}

/* hinit (common.rs:72-118): Hairer's starting step from f0 = f(x, y) and one
 * explicit Euler probe f1 = f(x + h0, y + h0 f0); order 5. */
static int hinit_step(workspace* w, const double* y, double x, double t1, double atol, double rtol,
                      double* f0, double* f1, double* yt, double* h_out) {
  const int dim = w->p->dim;
  int rc;
  if ((rc = rhs_eval(w, y, x, f0))) return rc;
  double posneg = sign_of(1.0, t1 - x), hmax = fabs(t1 - x), d0 = 0, d1 = 0, d2 = 0;
  for (int i = 0; i < dim; ++i) {
    double sci = atol + fabs(y[i]) * rtol;
    d0 += (y[i] / sci) * (y[i] / sci);
    d1 += (f0[i] / sci) * (f0[i] / sci);
  }
  double h0 = (d0 < 1.0e-10 || d1 < 1.0e-10) ? 1.0e-6 : 0.01 * sqrt(d0 / d1);
  h0 = fmin(h0, hmax);
  h0 = sign_of(h0, posneg);
  for (int i = 0; i < dim; ++i) yt[i] = y[i] + f0[i] * h0;
  if ((rc = rhs_eval(w, yt, x + h0, f1))) return rc;
  for (int i = 0; i < dim; ++i) {
    double sci = atol + fabs(y[i]) * rtol;
    d2 += ((f1[i] - f0[i]) / sci) * ((f1[i] - f0[i]) / sci);
  }
  d2 = sqrt(d2) / h0;
  double h1 = fmax(sqrt(d1), fabs(d2)) <= 1.0e-15 ? fmax(1.0e-6, fabs(h0) * 1.0e-3)
                                                  : pow(0.01 / fmax(sqrt(d1), d2), 1.0 / 5.0);
  *h_out = sign_of(fmin(100.0 * fabs(h0), fmin(h1, hmax)), posneg);
  return 0;
}

int lbo_hinit(lbo_plan* plan, const double* y, double t0, double t1, double abstol, double reltol,
              int32_t n_override, const int32_t* override_slots, const double* override_values,
              double* h_out) {
  const int dim = plan->dim;
  double* buf = calloc((size_t)dim * 3, sizeof(double));
  workspace w;
  ws_init(&w, plan);
  ws_set_overrides(&w, n_override, override_slots, (const cpx*)override_values);
  int rc = hinit_step(&w, y, t0, t1, abstol, reltol, buf, buf + dim, buf + 2 * dim, h_out);
  ws_free(&w);
  free(buf);
  return rc;
}

int lbo_eval_expression(int32_t n_instr, const int32_t* op, const int32_t* index, const int32_t* argc,
                        const int32_t* function, const double* re, const double* im, int32_t n_slots,
                        const double* slots, double t, double* out) {
  if (n_instr <= 0) return fail("empty expression");
  instr_t* ins = calloc((size_t)n_instr, sizeof(instr_t));
  for (int32_t i = 0; i < n_instr; ++i) {
    if (op[i] == OP_SLOT && (index[i] < 0 || index[i] >= n_slots)) {
      free(ins);
      return fail("slot index %d out of range", index[i]);
    }
    ins[i] = (instr_t){op[i], index[i], argc[i], function[i], re[i], im[i]};
  }
  expr_t e = {0, n_instr, 0};
  cpx v;
  int rc = eval_expr(ins, &e, (const cpx*)slots, t, &v);
  free(ins);
  if (!rc) {
    out[0] = v.re;
    out[1] = v.im;
  }
  return rc;
}

static int solve_one(const lbo_plan* p, workspace* w, const lbo_solve_args* a, const double* y0,
                     double t0, double t1, collector* out, ode_stats* st) {
  int dim = p->dim;
  memset(st, 0, sizeof(*st));
  if (a->solver >= LB_SOLVER_FIXED_DOPRI5) return solve_fixed(p, w, a, y0, t0, t1, out, st);
  if (a->maxiters <= 0) return fail("maxiters must be positive");
  if (t1 < t0) { st->status = LB_TRAJ_BAD_SPAN; st->fail_x = t0; return fail("only forward integration"); }
  if (t1 == t0) {
    if (a->save_start) coll_push(out, t0, y0);
    return 0;
  }
  /* save plan (common.rs:124-151) */
  double* sp = NULL;
  int64_t nsp = 0;
  int has_sp = a->saveat != NULL || a->saveat_is_t1;
  if (a->saveat_is_t1) {
    /* output_when="final": the save plan {t1} of this trajectory (python_api.rs:589-617) */
    sp = malloc(sizeof(double));
    sp[nsp++] = t1;
  } else if (has_sp) {
    sp = malloc(sizeof(double) * (a->n_saveat + 1));
    if (a->save_start) sp[nsp++] = t0;
    for (int64_t q = 0; q < a->n_saveat; ++q) {
      double tt = a->saveat[q];
      if (a->save_start && close_to_start(tt, t0)) continue;
      if (!isfinite(tt)) { free(sp); return fail("saveat contains non-finite value"); }
      if (tt < t0 - time_tol(t0) || tt > t1 + time_tol(t1)) {
        free(sp);
        return fail("saveat value %g is outside [%g, %g]", tt, t0, t1);
      }
      sp[nsp++] = tt;
    }
  }
  tableau tb = a->solver == LB_SOLVER_TSIT5 ? tsit5_tab() : dopri5_tab();
  double* buf = calloc((size_t)dim * 17, sizeof(double));
  instr_t* ev_ins = event_program(a);
  double* ye_buf = buf + 16 * dim; /* the state at an event's root */
  double* y = buf;
  double* yn = buf + dim;
  double* yt = buf + 2 * dim;
  double* ys = buf + 3 * dim;
  double* ds = buf + 4 * dim;
  double* f0 = buf + 5 * dim;
  double* f1 = buf + 6 * dim;
  double* k[8];
  for (int s = 1; s <= 7; ++s) k[s] = buf + (7 + s) * dim;
  memcpy(y, y0, sizeof(double) * dim);
  int rc = 0;
  double x = t0, h;
  if (isfinite(a->dt) && a->dt > 0.0) {
    h = fmin(a->dt, t1 - t0);
  } else {
    if ((rc = hinit_step(w, y, x, t1, a->abstol, a->reltol, f0, f1, yt, &h))) goto done;
  }
  controller ctrl;
  ctrl_init(&ctrl, t0, t1);
  double pn = sign_of(1.0, t1 - t0);
  if ((rc = rhs_eval(w, y, x, k[1]))) goto done;
  st->rhs_calls += 1;
  double ev_old = 0.0;
  if (a->event_kind && (rc = event_value(w, a, ev_ins, x, y, &ev_old))) {
    if (rc == EVENT_COMPLEX) { st->status = LB_TRAJ_EVENT_COMPLEX; st->fail_x = x; }
    goto done;
  }
  int64_t si = 0;
  if (has_sp) {
    while (si < nsp && close_to_start(sp[si], t0)) coll_push(out, sp[si++], y);
  } else if (a->save_start) {
    coll_push(out, x, y);
  }
  if (a->integral_out && a->integral_save_start) side_push(out, x, y);
  int64_t ns = 0, nst = 0, ia = 0;
  int last = 0;
  while (!last) {
    if (ns > a->maxiters) {
      st->status = LB_TRAJ_MAXITERS; st->fail_x = x;
      rc = fail("Stopped at x=%.17g. Need more than %lld steps.", x, (long long)ns);
      goto done;
    }
    if (0.1 * fabs(h) <= DBL_EPSILON * fabs(x)) {
      st->status = LB_TRAJ_UNDERFLOW; st->fail_x = x;
      rc = fail("Stopped at x=%.17g. Step size underflow.", x);
      goto done;
    }
    if ((x + 1.01 * h - t1) * pn > 0.0) {
      h = t1 - x;
      last = 1;
    }
    ns += 1;
    for (int s = 2; s <= 6; ++s) {
      fill_stage(&tb, s, dim, y, h, k, yt);
      if (s == 6) memcpy(ys, yt, sizeof(double) * dim);
      if ((rc = rhs_eval(w, yt, x + h * tb.c[s], k[s]))) goto done;
    }
    /* solution: all six A7j terms summed (A72 is 0 for Dopri5 but present in its sum) */
    for (int i = 0; i < dim; ++i) {
      double acc = tb.a[7][1] * k[1][i];
      for (int j = 2; j <= 6; ++j) acc = acc + tb.a[7][j] * k[j][i];
      yn[i] = y[i] + h * acc;
    }
    if ((rc = rhs_eval(w, yn, x + h * tb.c[7], k[7]))) goto done;
    st->rhs_calls += 6;
    double err = 0.0;
    for (int i = 0; i < dim; ++i) {
      double sc = a->abstol + fmax(fabs(y[i]), fabs(yn[i])) * a->reltol;
      double acc;
      if (tb.is_tsit5) {
        acc = tb.e[1] * k[1][i];
        for (int j = 2; j <= 7; ++j) acc = acc + tb.e[j] * k[j][i];
      } else {
        acc = tb.e[1] * k[1][i] + tb.e[3] * k[3][i] + tb.e[4] * k[4][i] + tb.e[5] * k[5][i] +
              tb.e[6] * k[6][i] + tb.e[7] * k[7][i];
      }
      double ei = h * acc;
      err += (ei / sc) * (ei / sc);
    }
    err = sqrt(err / (double)dim);
    double hn;
    int acc = ctrl_accept(&ctrl, err, h, &hn);
    if (acc) {
      st->accepted += 1;
      if (st->accepted % 1000 == 0 || ia > 0) {
        double n2 = 0.0, d2 = 0.0;
        for (int i = 0; i < dim; ++i) {
          double dk = k[7][i] - k[6][i], dy = yn[i] - ys[i];
          n2 += dk * dk;
          d2 += dy * dy;
        }
        double hl = d2 > 0.0 ? h * sqrt(n2 / d2) : 0.0;
        if (hl > 3.25) {
          ia += 1;
          nst = 0;
          if (ia == 15) {
            st->status = LB_TRAJ_STIFF; st->fail_x = x;
            rc = fail("Stiff at x=%.17g.", x);
            goto done;
          }
        } else {
          nst += 1;
          if (nst == 6) ia = 0;
        }
      }
      double xo = x;
      x += h;
      /* stop event (dopri5.rs:280-317, tsit5.rs:274-309): a root by 60
         bisections on the dense output of the step */
      int ev_hit = 0;
      double te = x;
      const double* ye = yn;
      if (a->event_kind) {
        double g1;
        if ((rc = event_value(w, a, ev_ins, x, yn, &g1))) {
          if (rc == EVENT_COMPLEX) { st->status = LB_TRAJ_EVENT_COMPLEX; st->fail_x = x; }
          goto done;
        }
        if (crossed_zero(ev_old, g1)) {
          ev_hit = 1;
          if (ev_old == 0.0) {
            te = xo;
            ye = y;
          } else if (g1 == 0.0) {
            te = x;
            ye = yn;
          } else {
            double lo = 0.0, hi = 1.0, glo = ev_old;
            for (int it = 0; it < 60; ++it) {
              const double mid = 0.5 * (lo + hi);
              if (tb.is_tsit5) dense_tsit5(mid, h, dim, y, k, ds);
              else dense_dopri5(mid, h, dim, y, yn, k, ds);
              double gm;
              if ((rc = event_value(w, a, ev_ins, xo + mid * h, ds, &gm))) {
                if (rc == EVENT_COMPLEX) { st->status = LB_TRAJ_EVENT_COMPLEX; st->fail_x = xo + mid * h; }
                goto done;
              }
              if (gm == 0.0) { lo = mid; hi = mid; break; }
              if ((glo < 0.0 && gm > 0.0) || (glo > 0.0 && gm < 0.0)) {
                hi = mid;
              } else {
                lo = mid;
                glo = gm;
              }
            }
            const double th = 0.5 * (lo + hi);
            if (tb.is_tsit5) dense_tsit5(th, h, dim, y, k, ye_buf);
            else dense_dopri5(th, h, dim, y, yn, k, ye_buf);
            te = xo + th * h;
            ye = ye_buf;
          }
        }
        ev_old = g1;
      }
      if (has_sp) {
        int dr = 0;
        const double tb_end = ev_hit ? te : x;
        while (si < nsp && sp[si] <= tb_end + time_tol(tb_end)) {
          double ts = sp[si];
          if (ts >= xo - time_tol(xo)) {
            if (fabs(ts - xo) <= time_tol(xo)) coll_push(out, ts, y);
            else if (fabs(ts - x) <= time_tol(x)) coll_push(out, ts, yn);
            else {
              if (tb.is_tsit5) dense_tsit5((ts - xo) / h, h, dim, y, k, ds);
              else dense_dopri5((ts - xo) / h, h, dim, y, yn, k, ds);
              dr = 1;
              coll_push(out, ts, ds);
            }
          }
          si++;
        }
        (void)dr;
      } else if (!ev_hit) {
        coll_push(out, x, yn);
      }
      side_push(out, x, yn);
      if (ev_hit) {
        /* push_unique (dopri5.rs:149-158, 345-351) */
        if (!(isfinite(out->last_push_t) && fabs(out->last_push_t - te) <= time_tol(te))) coll_push(out, te, ye);
        st->event_fired = 1;
        st->event_t = te;
        break;
      }
      double* tmp = y; y = yn; yn = tmp;
      memcpy(k[1], k[7], sizeof(double) * dim);
      if (last) break;
    } else {
      last = 0;
      if (st->accepted >= 1) st->rejected += 1;
    }
    h = hn;
  }
  if (has_sp && si != nsp && !st->event_fired)
    rc = fail("ended before all saveat: %lld of %lld", (long long)si, (long long)nsp);
done:
  if (rc == LB_TRAJ_FIELD_BELOW || rc == LB_TRAJ_FIELD_ABOVE) st->status = rc;
  if (rc && st->status == LB_TRAJ_OK) st->status = 5;
  free(buf);
  free(sp);
  free(ev_ins);
  return rc;
}

/* ---------------- batch driver ------------------------------------------ */

typedef struct {
  const lbo_plan* p;
  const lbo_solve_args* a;
  int64_t n_traj;
  const double* y0;
  int y0_shared;
  const double* t1_per_traj;
  int n_override;
  const int32_t* ov_slots;
  const double* ov_values;
  int64_t capacity, width;
  double *times, *values, *fail_x;
  int64_t *n_times, *stats;
  int64_t next;
  int first_error_traj;
  char first_error[512];
  pthread_mutex_t mu;
} batch_ctx;

static void* batch_worker(void* arg) {
  batch_ctx* b = arg;
  workspace w;
  ws_init(&w, b->p);
  for (;;) {
    int64_t i = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
    if (i >= b->n_traj) break;
    const double* y0 = b->y0 + (b->y0_shared ? 0 : i * (int64_t)b->p->dim);
    ws_set_overrides(&w, b->n_override, b->ov_slots,
                     b->n_override ? (const cpx*)(b->ov_values + 2 * i * b->n_override) : NULL);
    collector c;
    memset(&c, 0, sizeof(c));
    c.a = b->a;
    c.p = b->p;
    c.capacity = b->capacity;
    c.width = b->width;
    c.times = b->times + i * b->capacity;
    c.values = b->values + i * b->capacity * b->width;
    c.last_t = NAN;
    c.last_push_t = NAN;
    c.side_last_t = NAN;
    ode_stats st;
    double t1 = b->t1_per_traj ? b->t1_per_traj[i] : b->a->t1;
    int rc = solve_one(b->p, &w, b->a, y0, b->a->t0, t1, &c, &st);
    b->n_times[i] = c.count;
    b->stats[4 * i + 0] = st.accepted;
    b->stats[4 * i + 1] = st.rejected;
    b->stats[4 * i + 2] = st.rhs_calls;
    b->stats[4 * i + 3] = rc ? (st.status ? st.status : 5) : 0;
    b->fail_x[i] = st.fail_x;
    if (b->a->integral_out) b->a->integral_out[i] = rc ? NAN : c.side_integral;
    if (b->a->event_t) b->a->event_t[i] = (!rc && st.event_fired) ? st.event_t : NAN;
    if (rc) {
      pthread_mutex_lock(&b->mu);
      if (b->first_error_traj < 0 || i < b->first_error_traj) {
        b->first_error_traj = (int)i;
        snprintf(b->first_error, sizeof(b->first_error), "%s", g_err);
      }
      pthread_mutex_unlock(&b->mu);
    }
  }
  ws_free(&w);
  return NULL;
}

int lbo_solve_batch(const lbo_plan* plan, const lbo_solve_args* args, int64_t n_traj,
                    const double* y0, int32_t y0_shared, const double* t1_per_traj,
                    int32_t n_override, const int32_t* override_slots,
                    const double* override_values, int32_t threads, int64_t capacity,
                    double* times, double* values, int64_t* n_times, int64_t* stats,
                    double* fail_x) {
  batch_ctx b;
  memset(&b, 0, sizeof(b));
  b.p = plan;
  b.a = args;
  b.n_traj = n_traj;
  b.y0 = y0;
  b.y0_shared = y0_shared;
  b.t1_per_traj = t1_per_traj;
  b.n_override = n_override;
  b.ov_slots = override_slots;
  b.ov_values = override_values;
  b.capacity = capacity;
  b.width = lbo_output_width(plan, args);
  b.times = times;
  b.values = values;
  b.n_times = n_times;
  b.stats = stats;
  b.fail_x = fail_x;
  b.first_error_traj = -1;
  pthread_mutex_init(&b.mu, NULL);
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (threads > n_traj) threads = (int)n_traj;
  if (threads < 1) threads = 1;
  pthread_t* tid = calloc(threads, sizeof(pthread_t));
  for (int i = 1; i < threads; ++i) pthread_create(&tid[i], NULL, batch_worker, &b);
  batch_worker(&b);
  for (int i = 1; i < threads; ++i) pthread_join(tid[i], NULL);
  free(tid);
  pthread_mutex_destroy(&b.mu);
  if (b.first_error_traj >= 0) {
    snprintf(g_err, sizeof(g_err), "trajectory %d: %s", b.first_error_traj, b.first_error);
    return LB_ERR_INVALID;
  }
  return LB_OK;
}
