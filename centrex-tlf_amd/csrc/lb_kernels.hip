// HIP kernels of the MI355X Lindblad solver (gfx950, FP64).
//
//   lb_solve_kernel<R, SOLVER>  persistent fused Dopri5/Tsit5 integrator: one
//                               trajectory per workgroup at a time, taken from a
//                               device work queue; stage evaluation, RHS, error
//                               norm, PI controller, FSAL and the photon-count
//                               trapezoid all stay on the device.
//   lb_rhs_kernel<R>            one RHS evaluation per workgroup (plan checks).
//   lb_queue_reset_kernel       clears the work-queue counter (graph-safe, no memset).
//
// Method (reference files): expanded sparse RHS, rust/src/lindblad/rhs.rs:633-695,
// 815-876; controller rust/src/ode/common.rs:32-70; Dopri5 rust/src/ode/dopri5.rs;
// Tsit5 rust/src/ode/tsit5.rs; trapezoid collector rust/src/ode/output.rs:315-330.
// Data layout: lb_device.h.

#include "lb_device.h"
#include "lb_variants.h"

#ifndef __HIPCC_RTC__
#include <float.h>
#else
#define DBL_EPSILON 2.220446049250313080847e-16
#define DBL_MAX 1.79769313486231570815e+308
#define DBL_MIN 2.22507385850720138309e-308
#define M_PI 3.14159265358979323846
#endif

namespace {

struct dcpx {
  double re, im;
};

// Read-only plan tables indexed by wave-uniform values are read through the
// constant address space, so the loads go through the scalar cache instead
// of a vector load plus readfirstlane.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cst(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

__device__ __forceinline__ dcpx mk(double r, double i) { return dcpx{r, i}; }
__device__ __forceinline__ dcpx cadd(dcpx a, dcpx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ dcpx csub(dcpx a, dcpx b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ dcpx cmul(dcpx a, dcpx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ dcpx cdiv(dcpx a, dcpx b) {
  double ns = b.re * b.re + b.im * b.im;
  return mk((a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns);
}
__device__ __forceinline__ dcpx from_polar(double r, double th) {
  double s, c;
  sincos(th, &s, &c);
  return mk(r * c, r * s);
}
// The transcendental leaves of the expression machine are kept out of line:
// inlined, their temporaries would set the register allocation of the whole
// integrator kernel.  They run a few times per step in at most six lanes.
//
// exp(z) = e^re (cos im, sin im); exact for a real argument.
__device__ __noinline__ dcpx dev_cexp(dcpx z) {
  if (z.im == 0.0) return mk(exp(z.re), 0.0);
  return from_polar(exp(z.re), z.im);
}
// x^y = exp(y ln x) in polar form (x^0 = 1, 0^y = 0).  A real exponent skips the
// exp/log factors, which are exactly 1 and 0 in that case; a positive real base
// skips the angle, which is exactly 0.
__device__ __noinline__ dcpx dev_cpow(dcpx a, dcpx b) {
  if (b.re == 0.0 && b.im == 0.0) return mk(1.0, 0.0);
  double r = (a.im == 0.0) ? fabs(a.re) : hypot(a.re, a.im);
  if (r == 0.0) return mk(0.0, 0.0);
  if (b.im == 0.0) {
    double m = (b.re == 2.0) ? r * r : pow(r, b.re);
    if (a.im == 0.0 && a.re > 0.0) return mk(m, 0.0);
    return from_polar(m, b.re * atan2(a.im, a.re));
  }
  double th = atan2(a.im, a.re);
  return from_polar(pow(r, b.re) * exp(-b.im * th), b.re * th + b.im * log(r));
}

__device__ double rem_euclid(double x, double m) {
  double r = fmod(x, m);
  return r < 0.0 ? r + fabs(m) : r;
}

__device__ double gaussian_2d(double x, double y, double a, double mx, double my, double sx,
                              double sy) {
  double dx = x - mx, dy = y - my;
  return a * exp(-((dx * dx) / (2.0 * sx * sx) + (dy * dy) / (2.0 * sy * sy)));
}

__device__ double rabi_from_intensity(double intensity, double coupling, double dipole) {
  const double hbar = 1.0545718176461565e-34, c = 299792458.0, eps0 = 8.8541878128e-12;
  return sqrt(intensity * 2.0 / (c * eps0)) * coupling * dipole / hbar;
}

// Tuple helpers.  The host passes a tuple argument as its L elements, one
// after the other (a tuple's structure is resolved when the plan is built),
// and the real parts are used, as the reference's as_real does for real input.

// Sum of L Gaussian spots at (x, y): a = x, y, amplitudes[L], xlocs[L],
// ylocs[L], sigma_x, sigma_y (rust/src/lindblad/eval.rs:250-275).
__device__ double multipass_intensity(const dcpx* a, int L) {
  const double x = a[0].re, y = a[1].re, sx = a[2 + 3 * L].re, sy = a[3 + 3 * L].re;
  double total = 0.0;
  for (int k = 0; k < L; ++k) total += gaussian_2d(x, y, a[2 + k].re, a[2 + L + k].re, a[2 + 2 * L + k].re, sx, sy);
  return total;
}

// Piecewise-linear interpolation, flat outside the grid; a = x, grid[L],
// values[L] (eval.rs:360-388).  A non-increasing interval reached by the scan
// gives NaN (the reference raises; constant grids are checked on the host).
__device__ double linear_interp(const dcpx* a, int L) {
  const double x = a[0].re;
  const dcpx* g = a + 1;
  const dcpx* v = a + 1 + L;
  if (L == 1 || x <= g[0].re) return v[0].re;
  for (int i = 0; i + 1 < L; ++i) {
    const double left = g[i].re, right = g[i + 1].re;
    if (right <= left) return __builtin_nan("");
    if (x <= right) {
      const double frac = (x - left) / (right - left);
      return v[i].re + frac * (v[i + 1].re - v[i].re);
    }
  }
  return v[L - 1].re;
}

// PCHIP slope at knot k of a strictly increasing grid: one-sided secants at
// the ends, the weighted harmonic mean of the neighbouring secants inside,
// zero at a local extremum (eval.rs:318-331, plan.rs:54-66).
__device__ double pchip_slope(const dcpx* g, const dcpx* v, int L, int k) {
  auto secant = [&](int i) { return (v[i + 1].re - v[i].re) / (g[i + 1].re - g[i].re); };
  if (k == 0) return secant(0);
  if (k == L - 1) return secant(L - 2);
  const double d0 = secant(k - 1), d1 = secant(k);
  const double h0 = g[k].re - g[k - 1].re, h1 = g[k + 1].re - g[k].re;
  if ((d0 > 0.0) != (d1 > 0.0) || (d0 < 0.0) != (d1 < 0.0) || d0 == 0.0 || d1 == 0.0) return 0.0;
  const double w1 = 2.0 * h1 + h0, w2 = h1 + 2.0 * h0;
  return (w1 + w2) / (w1 / d0 + w2 / d1);
}

// pchip_interp in the Hermite form of eval.rs:290-351 (constant outside the
// grid), a = x, grid[L], values[L].  A non-increasing grid gives NaN.
__device__ double pchip_hermite(const dcpx* a, int L) {
  const double x = a[0].re;
  const dcpx* g = a + 1;
  const dcpx* v = a + 1 + L;
  if (L == 1) return v[0].re;
  for (int i = 0; i + 1 < L; ++i)
    if (g[i + 1].re - g[i].re <= 0.0) return __builtin_nan("");
  if (x <= g[0].re) return v[0].re;
  if (x >= g[L - 1].re) return v[L - 1].re;
  int i = 0;
  while (i < L - 2 && !(x < g[i + 1].re)) ++i;
  const double h = g[i + 1].re - g[i].re;
  const double t = (x - g[i].re) / h, t2 = t * t, t3 = t2 * t;
  const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
  const double h01 = -2.0 * t3 + 3.0 * t2, h11 = t3 - t2;
  return h00 * v[i].re + h10 * h * pchip_slope(g, v, L, i) + h01 * v[i + 1].re +
         h11 * h * pchip_slope(g, v, L, i + 1);
}

// pchip_interp through the reference's precomputed table (plan.rs:33-115):
// the cubic of interval i in Horner form, c0 + dx (c1 + dx (c2 + dx c3)),
// the interval the last knot at or below x (the last one at or beyond the
// grid's end), the first value at or below its start.  a = x, grid[L],
// values[L] of the table.
__device__ double pchip_table(const dcpx* a, int L) {
  const double x = a[0].re;
  const dcpx* g = a + 1;
  const dcpx* v = a + 1 + L;
  if (x <= g[0].re) return v[0].re;
  int i;
  double dx;
  if (x >= g[L - 1].re) {
    i = L - 2;
    dx = g[L - 1].re - g[L - 2].re;
  } else {
    i = 0;
    while (i < L - 2 && g[i + 1].re <= x) ++i;
    dx = x - g[i].re;
  }
  const double h = g[i + 1].re - g[i].re, delta = (v[i + 1].re - v[i].re) / h;
  const double d0 = pchip_slope(g, v, L, i), d1 = pchip_slope(g, v, L, i + 1);
  const double c2 = (3.0 * delta - 2.0 * d0 - d1) / h, c3 = (d0 + d1 - 2.0 * delta) / (h * h);
  return v[i].re + dx * (d0 + dx * (c2 + dx * c3));
}

// Scalar helpers of rust/src/lindblad/eval.rs:155-451, 664-789 (ids 1-7, 9,
// 11, 12, 13, 15), the tuple helpers (8, 10, 14, 16) and the table form of
// pchip_interp (17).
__device__ __noinline__ dcpx dev_helper(int id, const dcpx* a, int argc) {
  const double ED_XTB = 0.315 * 1.602176634e-19 * 5.29177210544e-11;
#define A(i) (a[i].re)
  switch (id) {
    case 1: return mk(gaussian_2d(A(0), A(1), A(2), A(3), A(4), A(5), A(6)), 0.0);
    case 2: {
      double th = A(7), ct = cos(th), st = sin(th), sx = A(5), sy = A(6);
      double p = ct * ct / (2.0 * sx * sx) + st * st / (2.0 * sy * sy);
      double q = sin(2.0 * th) / (2.0 * sx * sx) - sin(2.0 * th) / (2.0 * sy * sy);
      double u = st * st / (2.0 * sx * sx) + ct * ct / (2.0 * sy * sy);
      double dx = A(0) - A(3), dy = A(1) - A(4);
      return mk(A(2) * exp(-(p * dx * dx + q * dx * dy + u * dy * dy)), 0.0);
    }
    case 3: {
      double phi = A(1) * sin(A(2) * A(0));
      return mk(cos(phi), sin(phi));
    }
    case 4: return mk(0.5 * (1.0 + (sin(A(1) * A(0) + A(2)) >= 0.0 ? 1.0 : -1.0)), 0.0);
    case 5: {
      double th = 0.5 * A(1) * sin(A(2) * A(0));
      double v = 0.5 * (cos(th) + sin(th));
      return mk(v, v);
    }
    case 6: return mk(rem_euclid((A(1) * A(0) + A(2)) / (2.0 * M_PI), 1.0), 0.0);
    case 7: {
      double period = A(1) + A(2), duty = A(1) / period;
      double frac = rem_euclid(2.0 * M_PI * A(0) / period + A(3), 2.0 * M_PI) / (2.0 * M_PI);
      return mk(frac < duty ? 1.0 : 0.0, 0.0);
    }
    case 9: return mk(rabi_from_intensity(A(0), A(1), argc > 2 ? A(2) : ED_XTB), 0.0);
    case 11: {
      double inten = gaussian_2d(A(0), A(1), A(2), A(3), A(4), A(5), A(6));
      return mk(rabi_from_intensity(inten, A(7), argc > 8 ? A(8) : ED_XTB), 0.0);
    }
    case 12: {
      double frac = rem_euclid(A(0) * A(2) + A(3) / (2.0 * M_PI), 1.0);
      if (frac <= 0.0) frac += 1.0;
      return mk(frac < A(1) ? 1.0 : 0.0, 0.0);
    }
    case 13: {
      long long k = (long long)floor((A(0) - A(1)) / A(2));
      return mk(k % 2 == 0 ? 1.0 : -1.0, 0.0);
    }
    case 15: {
      double dx = A(0) - A(1);
      return mk(exp(-dx * dx / (2.0 * A(2) * A(2))), 0.0);
    }
    case 8: return mk(multipass_intensity(a, (argc - 4) / 3), 0.0);
    case 10: {
      const int L = (argc - 5) % 3 == 0 ? (argc - 5) / 3 : (argc - 6) / 3;
      return mk(rabi_from_intensity(multipass_intensity(a, L), A(4 + 3 * L), argc == 6 + 3 * L ? A(5 + 3 * L) : ED_XTB),
                0.0);
    }
    case 14: return mk(linear_interp(a, (argc - 1) / 2), 0.0);
    case 16: return mk(pchip_hermite(a, (argc - 1) / 2), 0.0);
    case 17: return mk(pchip_table(a, (argc - 1) / 2), 0.0);
    default: return mk(__builtin_nan(""), 0.0);  // rejected at plan creation
  }
#undef A
}

// Stack machine of rust/src/lindblad/eval.rs:791-911 (scalar form).  The plan
// builder has checked ops, helper ids and the stack depth.
// Builtins 1-7 of rust/src/lindblad/eval.rs:135-153 (num-complex formulas).
__device__ __noinline__ dcpx dev_builtin(int f, dcpx v) {
  switch (f) {
    case 1: return mk(sin(v.re) * cosh(v.im), cos(v.re) * sinh(v.im));
    case 2: return mk(cos(v.re) * cosh(v.im), -(sin(v.re) * sinh(v.im)));
    case 3: {
      double tr = v.re + v.re, ti = v.im + v.im, d = cos(tr) + cosh(ti);
      return mk(sin(tr) / d, sinh(ti) / d);
    }
    case 4: return dev_cexp(v);
    case 5: return mk(hypot(v.re, v.im), 0.0);
    case 6: return mk(v.re, 0.0);
    default: return mk(v.im, 0.0);
  }
}

// x^c for a real constant c, with cs = cos(c*pi) and sn = sin(c*pi) from the
// host: the polar form of dev_cpow with the angle of a negative real base
// (atan2(+-0, x<0) = +-pi) folded into the constants.
__device__ __forceinline__ dcpx dev_powc(dcpx a, double c, double cs, double sn) {
  if (c == 0.0) return mk(1.0, 0.0);
  if (a.im != 0.0) return dev_cpow(a, mk(c, 0.0));
  const double r = fabs(a.re);
  if (r == 0.0) return mk(0.0, 0.0);
  const double m = (c == 2.0) ? r * r : pow(r, c);
  if (a.re > 0.0) return mk(m, 0.0);
  return mk(m * cs, signbit(a.im) ? -(m * sn) : m * sn);
}

// Stack machine of rust/src/lindblad/eval.rs:791-911 (scalar form).  The plan
// builder has checked ops, helper ids and the stack depth.  The top of the
// stack lives in registers and the entries below it in this lane's LDS
// window `stk` (nothing goes to scratch memory).  Each 32-byte instruction
// record comes in with one scalar load, issued one instruction ahead.
__device__ __forceinline__ DevOp fetch_op(const __attribute__((address_space(4))) DevOp* ins, int p) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) u32x4 cu4;
  const cu4* w = reinterpret_cast<cu4*>(ins + p);
  const u32x4 a = w[0], b = w[1];
  DevOp q;
  q.code = (int32_t)a.x;
  q.index = (int32_t)a.y;
  q.re = __hiloint2double((int)a.w, (int)a.z);
  q.im = __hiloint2double((int)b.y, (int)b.x);
  q.aux = __hiloint2double((int)b.w, (int)b.z);
  return q;
}

__device__ __forceinline__ dcpx vm_eval(const __attribute__((address_space(4))) DevOp* ins, int start,
                                        int end, const dcpx* slots,
                                        double t, dcpx* stk) {
  dcpx top = mk(0.0, 0.0);
  if (start >= end) return top;
  int d = 0;  // entries on the stack; entry d-1 is `top`
  DevOp next = fetch_op(ins, start);
  for (int p = start; p < end; ++p) {
    const DevOp q = next;
    if (p + 1 < end) next = fetch_op(ins, p + 1);
    const int op = q.code & 0xFF;
    if (op <= 4) {
      if (d > 0) stk[d - 1] = top;
      ++d;
      top = op == 1 ? mk(q.re, q.im) : op == 2 ? slots[q.index] : mk(t, 0.0);
      continue;
    }
    const int argc = (q.code >> 8) & 0xFF, function = q.code >> 16;
    switch (op) {
      case 5: top = cadd(stk[d - 2], top); --d; break;
      case 6: top = csub(stk[d - 2], top); --d; break;
      case 7: top = cmul(stk[d - 2], top); --d; break;
      case 8: top = cdiv(stk[d - 2], top); --d; break;
      case 9: top = dev_cpow(stk[d - 2], top); --d; break;
      case 10: top = mk(-top.re, -top.im); break;
      case 11: top.im = -top.im; break;
      case 12: {
        const dcpx v = argc == 1 ? top : stk[d - argc];
        if (function == 4 && v.im == 0.0) top = mk(exp(v.re), 0.0);
        else top = dev_builtin(function, v);
        d -= argc - 1;
        break;
      }
      case 13:
        stk[d - 1] = top;
        top = dev_helper(function, &stk[d - argc], argc);
        d -= argc - 1;
        break;
      case LB_OP_POWC: top = dev_powc(top, q.re, q.im, q.aux); break;
      default: {
        const double rv = top.re, lv = stk[d - 2].re;
        --d;
        const bool res = op == 14 ? lv > rv : op == 15 ? lv >= rv : op == 16 ? lv < rv
                       : op == 17 ? lv <= rv : op == 18 ? lv == rv : lv != rv;
        top = mk(res ? 1.0 : 0.0, 0.0);
        break;
      }
    }
  }
  return top;
}

// Butcher tableaus (rust/src/ode/dopri5.rs:5-37, rust/src/ode/tsit5.rs:5-38):
// c(s) the node of stage s (c(1) = 0, c(7) = 1), a(s, j), the error weights e(j).
template <int SOLVER>
struct Tab;

template <>
struct Tab<0> {
  __device__ static constexpr double c(int s) {
    return s <= 1 ? 0.0 : s == 2 ? 1.0 / 5.0 : s == 3 ? 3.0 / 10.0 : s == 4 ? 4.0 / 5.0 : s == 5 ? 8.0 / 9.0 : 1.0;
  }
  __device__ static constexpr double a(int s, int j) {
    return s == 2 ? (j == 1 ? 1.0 / 5.0 : 0.0)
         : s == 3 ? (j == 1 ? 3.0 / 40.0 : j == 2 ? 9.0 / 40.0 : 0.0)
         : s == 4 ? (j == 1 ? 44.0 / 45.0 : j == 2 ? -56.0 / 15.0 : j == 3 ? 32.0 / 9.0 : 0.0)
         : s == 5 ? (j == 1 ? 19372.0 / 6561.0 : j == 2 ? -25360.0 / 2187.0 : j == 3 ? 64448.0 / 6561.0
                     : j == 4 ? -212.0 / 729.0 : 0.0)
         : s == 6 ? (j == 1 ? 9017.0 / 3168.0 : j == 2 ? -355.0 / 33.0 : j == 3 ? 46732.0 / 5247.0
                     : j == 4 ? 49.0 / 176.0 : j == 5 ? -5103.0 / 18656.0 : 0.0)
                  : (j == 1 ? 35.0 / 384.0 : j == 2 ? 0.0 : j == 3 ? 500.0 / 1113.0 : j == 4 ? 125.0 / 192.0
                     : j == 5 ? -2187.0 / 6784.0 : j == 6 ? 11.0 / 84.0 : 0.0);
  }
  __device__ static constexpr double e(int j) {
    return j == 1 ? 71.0 / 57600.0 : j == 2 ? 0.0 : j == 3 ? -71.0 / 16695.0 : j == 4 ? 71.0 / 1920.0
         : j == 5 ? -17253.0 / 339200.0 : j == 6 ? 22.0 / 525.0 : -1.0 / 40.0;
  }
};

template <>
struct Tab<1> {
  __device__ static constexpr double c(int s) {
    return s <= 1 ? 0.0 : s == 2 ? 0.161 : s == 3 ? 0.327 : s == 4 ? 0.9 : s == 5 ? 0.9800255409045097 : 1.0;
  }
  __device__ static constexpr double a(int s, int j) {
    return s == 2 ? (j == 1 ? 0.161 : 0.0)
         : s == 3 ? (j == 1 ? -0.008480655492356989 : j == 2 ? 0.335480655492357 : 0.0)
         : s == 4 ? (j == 1 ? 2.8971530571054935 : j == 2 ? -6.359448489975075 : j == 3 ? 4.3622954328695815 : 0.0)
         : s == 5 ? (j == 1 ? 5.325864828439257 : j == 2 ? -11.748883564062828 : j == 3 ? 7.4955393428898365
                     : j == 4 ? -0.09249506636175525 : 0.0)
         : s == 6 ? (j == 1 ? 5.86145544294642 : j == 2 ? -12.92096931784711 : j == 3 ? 8.159367898576159
                     : j == 4 ? -0.071584973281401 : j == 5 ? -0.028269050394068383 : 0.0)
                  : (j == 1 ? 0.09646076681806523 : j == 2 ? 0.01 : j == 3 ? 0.4798896504144996
                     : j == 4 ? 1.379008574103742 : j == 5 ? -3.290069515436081 : j == 6 ? 2.324710524099774 : 0.0);
  }
  __device__ static constexpr double e(int j) {
    return j == 1 ? -0.0017800110522257771 : j == 2 ? -0.0008164344596567469 : j == 3 ? 0.007880878010261995
         : j == 4 ? -0.1447110071732629 : j == 5 ? 0.5823571654525552 : j == 6 ? -0.458082105929187
         : 1.0 / 66.0;
  }
};

// Stage input y + h * sum_j a_sj k_j (dopri5.rs:45-86) for stage S; S = 7 is
// the solution.  Zero tableau entries are skipped at compile time.
template <int R, int SOLVER, int S>
__device__ __forceinline__ void stage_combo(double h, const double (&y_r)[R], const double (&y_i)[R],
                                            const double (&k_r)[8][R], const double (&k_i)[8][R],
                                            double (&o_r)[R], double (&o_i)[R]) {
  using T = Tab<SOLVER>;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double sr = T::a(S, 1) * k_r[1][r], si = T::a(S, 1) * k_i[1][r];
#pragma unroll
    for (int j = 2; j < S; ++j) {
      if (T::a(S, j) != 0.0) {
        sr = fma(T::a(S, j), k_r[j][r], sr);
        si = fma(T::a(S, j), k_i[j][r], si);
      }
    }
    o_r[r] = fma(h, sr, y_r[r]);
    o_i[r] = fma(h, si, y_i[r]);
  }
}

template <int R>
__device__ __forceinline__ void copy_k(const double (&a)[R], const double (&b)[R], double (&ka)[R],
                                       double (&kb)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    ka[r] = a[r];
    kb[r] = b[r];
  }
}

// Per-thread view of the slots it owns: LDS publication index and flags
// (bit 0 valid, bit 1 diagonal, bits 2.. diagonal state index).
template <int R>
struct Slots {
  int pub[R];
  int flags[R];
  __device__ __forceinline__ bool valid(int r) const { return flags[r] & 1; }
  __device__ __forceinline__ bool diag(int r) const { return flags[r] & 2; }
  __device__ __forceinline__ int state(int r) const { return flags[r] >> 2; }
};

// LDS carve-up of one workgroup (sizes from the plan, see lds_bytes_for).
struct Lds {
  double* pub;    // 2 buffers x 2*n_pub
  double* selfc;  // 2*S folded self coefficients
  double* wts;    // n integral weights per state
  dcpx* coef;     // 8 x (K+1): [0] t0 and hinit, [1] hinit second point, [2..7] stages
  dcpx* slots;    // 8 x n_slots
  double* red;    // 4 x nblk block sums
  dcpx* vm;       // 8 lanes x vm_depth expression stacks
  double2* park;  // 128 entries: single-wave state parked around the expression machine
  double2* gf;    // gather factors (when the plan stages its tables in LDS)
  uint32_t* gw;   // gather words
  int* misc;      // [0] queue value
  double* evv;    // n + 2: stop-event populations, [n] expression value, [n + 1] complex flag
};

__device__ __forceinline__ Lds carve(const DevPlan& P) {
  extern __shared__ double smem[];
  Lds L;
  L.pub = smem;
  L.selfc = L.pub + 4 * (size_t)P.n_pub;
  L.wts = L.selfc + 2 * (size_t)P.S;
  L.coef = reinterpret_cast<dcpx*>(L.wts + ((P.n + 1) & ~1));
  L.slots = L.coef + 8 * (P.K + 1);
  L.red = reinterpret_cast<double*>(L.slots + 8 * P.n_slots);
  L.vm = reinterpret_cast<dcpx*>(L.red + 4 * P.nblk);
  L.park = reinterpret_cast<double2*>(L.vm + 8 * P.vm_depth);
  L.gf = L.park + 128;
  const size_t cells = P.lds_tables ? (size_t)P.g_rows * 64 : 0;
  L.gw = reinterpret_cast<uint32_t*>(L.gf + cells);
  L.misc = reinterpret_cast<int*>(L.gw + ((cells + 3) & ~(size_t)3));
  L.evv = reinterpret_cast<double*>(L.misc + 4);
  return L;
}

// One gather term: acc += f * c * x_in (rhs.rs:815-876 term semantics), with
// the input element (x_r, x_raw_i) already fetched.
__device__ __forceinline__ void gather_apply(uint32_t w, double2 f, const dcpx* coef, double x_r,
                                             double x_raw_i, double& accr, double& acci) {
  // A static term uses its factor as it is.  The coefficient is loaded for
  // every term (entry 0 for a static one) and the factor selected after, so
  // no lane branches and the load issues with the others.
  const uint32_t cid = w >> 24;
  const bool stat = cid == LB_STATIC_COEF;
  dcpx c = coef[stat ? 0u : cid];
  c.im = (w & (1u << 22)) ? -c.im : c.im;
  const double pr = fma(f.x, c.re, -(f.y * c.im));
  const double pi = fma(f.x, c.im, f.y * c.re);
  const double vr = stat ? f.x : pr, vi = stat ? f.y : pi;
  double x_i = (w & (1u << 20)) ? 0.0 : x_raw_i;
  x_i = (w & (1u << 21)) ? -x_i : x_i;
  // explicit fused operations: the library is built without implicit
  // contraction, so every kernel variant rounds identically
  accr = fma(-vi, x_i, fma(vr, x_r, accr));
  acci = fma(vi, x_r, fma(vr, x_i, acci));
}

// The same term with its input read from the LDS publication buffer.
__device__ __forceinline__ void gather_term(uint32_t w, double2 f, const dcpx* coef, const double* pubbuf,
                                            double& accr, double& acci) {
  const uint32_t p = w & 0xFFFFFu;
  gather_apply(w, f, coef, pubbuf[2 * p], pubbuf[2 * p + 1], accr, acci);
}

// Single-wave form for plans of at most 64 elements: lane = slot, the gather
// inputs come from the owning lanes by cross-lane permutes instead of a
// publish / barrier / LDS-read round trip.  The publication index of a plan
// this small is its slot, so the table words are shared with the multi-wave
// form, and the arithmetic is the same.  All rows' table entries and inputs
// are fetched before the first term is applied.
template <int G>
__device__ __forceinline__ void rhs_wave(const Lds& L, int rows, int lane, double2 selfv, bool diag, double xr,
                                         double xi, double& kr, double& ki, const dcpx* coef) {
  double accr = fma(selfv.x, xr, -(selfv.y * xi));
  double acci = fma(selfv.x, xi, selfv.y * xr);
  uint32_t w[G];
  double2 f[G];
  double sr[G], si[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if (j < rows) {
      w[j] = L.gw[j * 64 + lane];
      f[j] = L.gf[j * 64 + lane];
    }
  }
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if (j < rows) {
      const int src = (int)(w[j] & 63u);
      sr[j] = __shfl(xr, src, 64);
      si[j] = __shfl(xi, src, 64);
    }
  }
#pragma unroll
  for (int j = 0; j < G; ++j)
    if (j < rows) gather_apply(w[j], f[j], coef, sr[j], si[j], accr, acci);
  kr = accr;
  ki = diag ? 0.0 : acci;
}

// The gather rows of one block, in row order.  The tables are read through
// generic pointers (LDS copy or global memory, one code path) with the next
// row's word and factor loaded before the current row is applied.
__device__ __forceinline__ void gather_block(const uint32_t* gw, const double2* gf, int row0, int rows,
                                             int lane, const dcpx* coef, const double* pubbuf,
                                             double& accr, double& acci) {
  if (rows <= 0) return;
  size_t idx = (size_t)row0 * 64 + lane;
  uint32_t w = gw[idx];
  double2 f = gf[idx];
  for (int j = 1; j < rows; ++j) {
    idx += 64;
    const uint32_t wn = gw[idx];
    const double2 fn = gf[idx];
    gather_term(w, f, coef, pubbuf, accr, acci);
    w = wn;
    f = fn;
  }
  gather_term(w, f, coef, pubbuf, accr, acci);
}

// d(rho)/dt for the slots of this thread: publish the cross-coupled elements,
// barrier, then self term + gathers.  `pubbuf` alternates between two buffers
// so one barrier per evaluation suffices.
template <int R>
__device__ __forceinline__ void rhs_eval(const DevPlan& P, const Lds& L, const Slots<R>& S,
                                         const double (&xr)[R], const double (&xi)[R], double (&kr)[R],
                                         double (&ki)[R], double* pubbuf, const dcpx* coef, int tid,
                                         int W) {
  const int lane = tid & 63;
  // wave-uniform block index: lets the row counts live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t* gw = P.lds_tables ? static_cast<const uint32_t*>(L.gw) : P.g_word;
  const double2* gf = P.lds_tables ? static_cast<const double2*>(L.gf) : reinterpret_cast<const double2*>(P.g_f);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (S.pub[r] >= 0) {
      pubbuf[2 * S.pub[r]] = xr[r];
      pubbuf[2 * S.pub[r] + 1] = xi[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int blk = wave + (W >> 6) * r;
    if (blk >= P.nblk) { kr[r] = 0.0; ki[r] = 0.0; continue; }
    const int s = tid + W * r;
    const double2 a = reinterpret_cast<const double2*>(L.selfc)[s];
    double accr = fma(a.x, xr[r], -(a.y * xi[r]));
    double acci = fma(a.x, xi[r], a.y * xr[r]);
    gather_block(gw, gf, cst(P.blk_row0)[blk], cst(P.blk_rows)[blk], lane, coef, pubbuf, accr, acci);
    kr[r] = accr;
    ki[r] = S.diag(r) ? 0.0 : acci;
  }
}

// Copy the gather tables into LDS once per workgroup (plans whose tables fit).
__device__ __forceinline__ void stage_tables(const DevPlan& P, const Lds& L, int tid, int W) {
  if (!P.lds_tables) return;
  const int cells = P.g_rows * 64;
  const double2* gf = reinterpret_cast<const double2*>(P.g_f);
  for (int i = tid; i < cells; i += W) {
    L.gf[i] = gf[i];
    L.gw[i] = P.g_word[i];
  }
}

// Sum over the 64 lanes of a wave with a fixed association order: neighbour
// pairs (xor 1, xor 2) by DPP quad permutes, then the mirrored halves of each
// 8- and 16-lane row, then the four rows as (r0 + r1) + (r2 + r3).  Every
// lane returns the same value; all 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_d(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_d<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_d<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_d<0x141>(x);  // row_half_mirror
  x += dpp_d<0x140>(x);  // row_mirror
  return (lane_d(x, 0) + lane_d(x, 16)) + (lane_d(x, 32) + lane_d(x, 48));
}

// Store the block sum of slot row r as red[q*nblk + blk].
__device__ __forceinline__ void block_store(const DevPlan& P, int tid, int W, int r, double* red, int q,
                                            double value) {
  const int s = tid + W * r;
  const double sum = wave_sum(value);
  if ((tid & 63) == 0 && s < P.S) red[q * P.nblk + (s >> 6)] = sum;
}

// Total of the stored block sums (call after a barrier): lane l adds blocks
// l, l+64, ... in order, then the lanes are summed as in wave_sum.  The same
// in every thread and independent of the workgroup size.
__device__ __forceinline__ double block_total(const DevPlan& P, const double* red, int q) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
  for (int b = lane; b < P.nblk; b += 64) v += red[q * P.nblk + b];
  return wave_sum(v);
}

__device__ __forceinline__ double time_tol(double t) { return 10.0 * DBL_EPSILON * fmax(fabs(t), 1.0); }

// Evaluate the parameter graph + coefficients for time t into slot copy `sl`
// and coefficient row `cf` (compounds first, in dependency order, then the
// coefficients); only_dynamic runs the time-dependent expressions only.
// n_prog < 0 evaluates the solver's K coefficient programs, n_prog = P.KD
// also the derivative programs of a differentiable plan; program j lands in
// cf[j] for j < K and in cf[j + 1] after the constant entry cf[K].  Each
// expression record comes in with one scalar load, issued one ahead.
__device__ __forceinline__ DevExpr fetch_expr(const __attribute__((address_space(4))) DevExpr* ex, int i) {
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  const i32x4 a = reinterpret_cast<const __attribute__((address_space(4))) i32x4*>(ex)[i];
  DevExpr r;
  r.start = a.x;
  r.end = a.y;
  r.dst = a.z;
  r.pad = 0;
  return r;
}

__device__ __forceinline__ void eval_coefficients(const DevPlan& P, dcpx* sl, dcpx* cf, double t,
                                                  bool only_dynamic, dcpx* stk, int n_prog = -1) {
  const bool all_programs = n_prog >= 0 && n_prog > P.K;
  const __attribute__((address_space(4))) DevExpr* ex = cst(only_dynamic ? P.ex_dyn : P.ex_all);
  const int n = only_dynamic ? (all_programs ? P.n_dyn : P.n_dyn_fwd) : (all_programs ? P.n_all : P.n_all_fwd);
  if (n > 0) {
    DevExpr next = fetch_expr(ex, 0);
    for (int i = 0; i < n; ++i) {
      const DevExpr r = next;
      if (i + 1 < n) next = fetch_expr(ex, i + 1);
      const dcpx v = vm_eval(cst(P.instr), r.start, r.end, sl, t, stk);
      if (r.dst >= 0) sl[r.dst] = v;
      else cf[-r.dst - 1] = v;
    }
  }
  cf[P.K] = mk(1.0, 0.0);
}

template <int R>
__device__ __forceinline__ void load_slots(const DevPlan& P, Slots<R>& S, int tid, int W) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = tid + W * r;
    const bool in = s < P.S;
    const int up = in ? P.slot_up[s] : -1;
    const int im = in ? P.slot_im[s] : -1;
    const int re = in ? P.slot_re[s] : 0;
    S.flags[r] = (up >= 0 ? 1 : 0) | (im < 0 ? 2 : 0) | ((im < 0 && up >= 0 ? re : 0) << 2);
    S.pub[r] = in ? P.slot_pub[s] : -1;
  }
}

// Per-trajectory parameter slots, part 1: base values (overrides follow in the
// kernels, which read them straight from their argument structs).
__device__ __forceinline__ void init_slots(const DevPlan& P, dcpx* sl, int tid, int W) {
  for (int i = tid; i < P.n_slots; i += W)
    sl[i] = i < P.n_base ? mk(P.base[2 * i], P.base[2 * i + 1]) : mk(0.0, 0.0);
  __syncthreads();
}

// Part 2: base-slot overrides of trajectory `traj`, from a row of a batch or a
// grid point with the last axis fastest (rust/src/lindblad/ode_batch.rs:353-391).
__device__ __forceinline__ void apply_overrides(const DevCall& C, int64_t traj, dcpx* sl) {
  for (int j = 0; j < C.n_override; ++j) {
    double vr, vi;
    if (C.n_axes > 0) {
      // n_traj < 2^31 (checked by the host), so 32-bit division suffices
      const uint32_t q = (uint32_t)traj / (uint32_t)C.axis_stride[j];
      const int64_t ai = C.axis_offset[j] + q % (uint32_t)C.axis_len[j];
      vr = C.axis_values[2 * ai];
      vi = C.axis_values[2 * ai + 1];
    } else {
      vr = C.override_values[2 * (traj * C.n_override + j)];
      vi = C.override_values[2 * (traj * C.n_override + j) + 1];
    }
    sl[C.override_slot[j]] = mk(vr, vi);
  }
}

// Fold the time-independent self terms of every slot with this trajectory's
// coefficient values into LDS (rhs.rs:901-919 folds static terms the same way).
__device__ __forceinline__ void fold_self(const DevPlan& P, int tid, int W, const dcpx* coef,
                                          double* selfc) {
  for (int s = tid; s < P.S; s += W) {
    dcpx acc = mk(0.0, 0.0);
    for (int j = 0; j < P.max_self; ++j) {
      const size_t e = (size_t)s * P.max_self + j;
      const int id = P.self_coef[e];
      if (id == -2) break;
      const dcpx f = mk(P.self_f[2 * e], P.self_f[2 * e + 1]);
      dcpx c = id < 0 ? mk(1.0, 0.0) : coef[id];
      if (P.self_conj[e]) c.im = -c.im;
      acc = cadd(acc, cmul(f, c));
    }
    selfc[2 * s] = acc.re;
    selfc[2 * s + 1] = acc.im;
  }
}

// ---- terminal stop events (rust/src/lindblad/ode_impl.rs:134-153)

// The time-dependent compounds (parameter-graph slots) at time t into slot
// copy `sl`; coefficient records are skipped.  One lane.
__device__ __forceinline__ void eval_compounds(const DevPlan& P, dcpx* sl, double t, dcpx* stk) {
  const __attribute__((address_space(4))) DevExpr* ex = cst(P.ex_dyn);
  for (int i = 0; i < P.n_dyn_fwd; ++i) {
    const DevExpr r = fetch_expr(ex, i);
    if (r.dst >= 0) sl[r.dst] = vm_eval(cst(P.instr), r.start, r.end, sl, t, stk);
  }
}

// g(t, y): the listed populations summed in list order minus the threshold,
// or the real part of the event expression over the slots `sl` (which hold
// the parameter values at t).  Every lane returns the same value; `complex_`
// reports an expression with |Im| > 1e-12 (rhs.rs:1230-1236).  Called by
// every lane of the workgroup (it synchronises).
template <int R>
__device__ __forceinline__ double event_value(const DevPlan& P, const DevCall& C, const Lds& L, const Slots<R>& S, int tid, int W,
                              double t, const dcpx* sl, const double (&vr)[R], bool& complex_) {
  double g;
  if (C.event_kind == LB_EVENT_POPULATION) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (S.valid(r) && S.diag(r)) L.evv[S.state(r)] = vr[r];
    __syncthreads();
    double v = 0.0;
    for (int k = 0; k < C.n_event; ++k) v += L.evv[C.event_idx[k]];
    g = v - C.event_threshold;
    complex_ = false;
  } else {
    if (tid == 0) {
      const dcpx v = vm_eval(cst(C.event_ops), 0, C.n_event, sl, t, L.vm);
      L.evv[P.n] = v.re;
      L.evv[P.n + 1] = fabs(v.im) > 1e-12 ? 1.0 : 0.0;
    }
    __syncthreads();
    g = L.evv[P.n];
    complex_ = L.evv[P.n + 1] != 0.0;
  }
  __syncthreads();  // the next evaluation rewrites evv
  return g;
}

__device__ __forceinline__ bool crossed_zero(double g0, double g1) {
  return g0 == 0.0 || g1 == 0.0 || (g0 < 0.0 && g1 > 0.0) || (g0 > 0.0 && g1 < 0.0);
}

// Dense output at theta in [0, 1] of the accepted step [x, x + h]
// (rust/src/ode/dopri5.rs:38-43, 104-143: Hairer's continuous extension;
// rust/src/ode/tsit5.rs:39-60, 121-142: Tsitouras' interpolant).
template <int R, int SOLVER>
__device__ __forceinline__ void dense_point(double th, double h, const double (&y_r)[R], const double (&y_i)[R],
                                            const double (&yn_r)[R], const double (&yn_i)[R],
                                            const double (&k_r)[8][R], const double (&k_i)[8][R],
                                            double (&o_r)[R], double (&o_i)[R]) {
  if constexpr (SOLVER == 0) {
    const double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
    const double th1 = 1.0 - th;
    auto one = [&](double y, double yn, double k1, double k3, double k4, double k5, double k6, double k7) {
      const double d = yn - y;
      const double b = k1 * h - d;
      const double r3 = -k7 * h + d - b;
      const double r4 = h * (D1 * k1 + D3 * k3 + D4 * k4 + D5 * k5 + D6 * k6 + D7 * k7);
      return y + (d + (b + (r3 + r4 * th1) * th) * th1) * th;
    };
#pragma unroll
    for (int r = 0; r < R; ++r) {
      o_r[r] = one(y_r[r], yn_r[r], k_r[1][r], k_r[3][r], k_r[4][r], k_r[5][r], k_r[6][r], k_r[7][r]);
      o_i[r] = one(y_i[r], yn_i[r], k_i[1][r], k_i[3][r], k_i[4][r], k_i[5][r], k_i[6][r], k_i[7][r]);
    }
  } else {
    const double t2 = th * th, t3 = t2 * th, t4 = t3 * th;
    const double b1 = 1.0 * th + -2.763706197274826 * t2 + 2.9132554618219126 * t3 + -1.0530884977290216 * t4;
    const double b2 = 0.1317 * t2 + -0.2234 * t3 + 0.1017 * t4;
    const double b3 = 3.9302962368947516 * t2 + -5.941033872131505 * t3 + 2.490627285651253 * t4;
    const double b4 = -12.411077166933676 * t2 + 30.33818863028232 * t3 + -16.548102889244902 * t4;
    const double b5 = 37.50931341651104 * t2 + -88.1789048947664 * t3 + 47.37952196281928 * t4;
    const double b6 = -27.896526289197286 * t2 + 65.09189467479366 * t3 + -34.87065786149661 * t4;
    const double b7 = 1.5 * t2 + -4.0 * t3 + 2.5 * t4;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      o_r[r] = y_r[r] + h * (b1 * k_r[1][r] + b2 * k_r[2][r] + b3 * k_r[3][r] + b4 * k_r[4][r] +
                             b5 * k_r[5][r] + b6 * k_r[6][r] + b7 * k_r[7][r]);
      o_i[r] = y_i[r] + h * (b1 * k_i[1][r] + b2 * k_i[2][r] + b3 * k_i[3][r] + b4 * k_i[4][r] +
                             b5 * k_i[5][r] + b6 * k_i[6][r] + b7 * k_i[7][r]);
    }
  }
}

// One output row of a trace (populations / selected / full; the integral
// row is written by the caller).
template <int R>
__device__ __forceinline__ void write_row(const DevPlan& P, const DevCall& C, const Slots<R>& S, int tid, int W,
                                          int64_t row, const double (&vr)[R], const double (&vi)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!S.valid(r)) continue;
    const int s = tid + W * r;
    if (C.output_kind == LB_OUT_POPULATIONS) {
      if (S.diag(r)) C.out_values[row * P.n + S.state(r)] = vr[r];
    } else if (C.output_kind == LB_OUT_FULL) {
      C.out_values[row * P.dim + P.slot_re[s]] = vr[r];
      if (!S.diag(r)) C.out_values[row * P.dim + P.slot_im[s]] = vi[r];
    } else if (C.output_kind == LB_OUT_SELECTED) {
      const int up = P.slot_up[s];
      for (int q = 0; q < C.n_selected; ++q) {
        if (C.sel_upper[q] != up) continue;
        const double im = (C.sel_flags[q] & 1) ? 0.0 : ((C.sel_flags[q] & 2) ? -vi[r] : vi[r]);
        C.out_values[2 * (row * C.n_selected + q)] = vr[r];
        C.out_values[2 * (row * C.n_selected + q) + 1] = im;
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Persistent integrator.
// ---------------------------------------------------------------------------
// G > 0 selects the single-wave form (R = 1, 64 threads, G gather rows in
// registers); G = 0 is the general multi-wave form.
template <int R, int SOLVER, int MAXW, int WAVES, int G, bool EV = false>
__global__ __launch_bounds__(MAXW, WAVES) void lb_solve_kernel(const DevPlan P, const DevCall C) {
  static_assert(G == 0 || (R == 1 && MAXW == 64), "the single-wave form holds one slot per lane");
  static_assert(!EV || G == 0, "stop events run in the multi-wave form");
  using T = Tab<SOLVER>;
  const int tid = threadIdx.x, W = blockDim.x;
  const Lds L = carve(P);
  const int KK = P.K + 1;
  Slots<R> S;
  load_slots<R>(P, S, tid, W);
  stage_tables(P, L, tid, W);
  const int wrows = G > 0 ? cst(P.blk_rows)[0] : 0;
  double2 selfv = double2{0.0, 0.0};

  for (;;) {
    __syncthreads();
    if (tid == 0) L.misc[0] = atomicAdd(C.queue, 1);
    __syncthreads();
    const int64_t local = L.misc[0];
    if (local >= C.n_traj) break;
    // queue position -> trajectory (longest-first orders); entries outside
    // the batch are skipped, so a bad order cannot address past the arrays
    const int64_t traj = C.queue_order ? (int64_t)C.queue_order[local + C.traj_offset] : local + C.traj_offset;
    if (traj < 0 || traj >= C.n_total) continue;
    const double t0 = C.t0;
    const double t1 = C.t1_per_traj ? C.t1_per_traj[traj] : C.t1;

    // ---- parameters and coefficients at t0, self-term folds, weights
    init_slots(P, L.slots, tid, W);
    if (tid == 0) {
      apply_overrides(C, traj, L.slots);
      eval_coefficients(P, L.slots, L.coef, t0, false, L.vm);
    }
    for (int i = tid; i < P.n; i += W) {
      double w = 0.0;
      for (int q = 0; q < C.n_weights; ++q)
        if (C.weight_index[q] == i) w += C.weight_value[q];
      L.wts[i] = w;
      if constexpr (EV) L.evv[i] = 0.0;  // stop-event populations outside a restricted plan stay zero
    }
    __syncthreads();
    for (int i = tid; i < 7 * P.n_slots; i += W) L.slots[P.n_slots + i] = L.slots[i % P.n_slots];
    for (int i = tid; i < 7 * KK; i += W) L.coef[KK + i] = L.coef[i % KK];
    fold_self(P, tid, W, L.coef, L.selfc);
    if constexpr (G > 0) {
      __syncthreads();  // coefficient rows and self folds are read below
      selfv = reinterpret_cast<const double2*>(L.selfc)[tid];
    }

    // ---- state
    double y_r[R], y_i[R];
    const double* y0 = C.y0 + (C.y0_shared ? 0 : traj * (int64_t)P.dim);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int s = tid + W * r;
      y_r[r] = S.valid(r) ? y0[P.slot_re[s]] : 0.0;
      y_i[r] = (S.valid(r) && !S.diag(r)) ? y0[P.slot_im[s]] : 0.0;
    }
    // a restricted plan needs y0 to vanish outside its elements
    bool bad_support = false;
    if (P.restricted) {
      if (tid == 0) L.misc[1] = 0;
      __syncthreads();
      for (int k = tid; k < P.dim; k += W)
        if (!P.upper_active[P.packed_upper[k]] && y0[k] != 0.0) L.misc[1] = 1;
      __syncthreads();
      bad_support = L.misc[1] != 0;
    }

    int64_t accepted = 0, rejected = 0, rhs_calls = 0, attempts = 0;
    int status = 0;
    double fail_x = 0.0, integral = 0.0, last_t = __builtin_nan(""), last_v = 0.0;
    // stop event: g at the last accepted point, the root's time once it fired,
    // and the time of the last pushed point (the root is pushed unless it is
    // that point, dopri5.rs:149-158)
    double ev_old = 0.0, ev_t = __builtin_nan(""), last_pt = __builtin_nan("");
    bool ev_fired = false;
    // pushed points (rust/src/ode/output.rs): the trapezoid integral and the
    // trace rows.  Pushing at every accepted step folds the integrand into
    // the error-norm reduction; a save plan (saveat) reduces per point.
    int64_t n_pushed = 0, si = 0;
    const bool step_push = C.n_saveat == 0;
    const bool rate = C.output_kind == LB_OUT_RATE;
    auto push = [&](double tp, const double (&vr)[R], const double (&vi)[R], bool reduce_integrand) {
      double v = last_v;
      if constexpr (EV) last_pt = tp;
      if ((C.want_integral || rate) && reduce_integrand) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          block_store(P, tid, W, r, L.red, 1, S.diag(r) && S.valid(r) ? L.wts[S.state(r)] * vr[r] : 0.0);
        __syncthreads();
        v = block_total(P, L.red, 1);
        __syncthreads();  // the next point reuses the reduction slots
        if (C.want_integral) {
          if (isfinite(last_t)) integral += 0.5 * (last_v + v) * (tp - last_t);
          last_t = tp;
          last_v = v;
        }
      }
      if (C.trace) {
        if (n_pushed < C.capacity) {
          const int64_t row = traj * C.capacity + n_pushed;
          if (C.output_kind == LB_OUT_INTEGRAL) {
            if (tid == 0) C.out_values[row] = integral;
          } else if (rate) {
            if (tid == 0) C.out_values[row] = v;  // weighted rate (output.rs:337-384)
          } else {
            write_row<R>(P, C, S, tid, W, row, vr, vi);
          }
          if (tid == 0 && C.out_times) C.out_times[row] = tp;
        }
        ++n_pushed;
      }
    };
    double k_r[8][R], k_i[8][R];
    int buf = 0;
#define PUB (L.pub + buf * 2 * P.n_pub)
#define LB_RHS(xr, xi, kr, ki, cf)                                                         \
  do {                                                                                     \
    if constexpr (G > 0) {                                                                 \
      rhs_wave<(G > 0 ? G : 1)>(L, wrows, tid, selfv, S.diag(0), xr[0], xi[0], kr[0], ki[0], cf); \
    } else {                                                                               \
      rhs_eval<R>(P, L, S, xr, xi, kr, ki, PUB, cf, tid, W);                               \
      buf ^= 1;                                                                            \
    }                                                                                      \
  } while (0)

    if (bad_support) {
      status = LB_TRAJ_BAD_SUPPORT;
      fail_x = t0;
    } else if (t1 < t0) {
      status = LB_TRAJ_BAD_SPAN;
      fail_x = t0;
    } else if (t1 == t0) {
      if (C.save_start) push(t0, y_r, y_i, true);
    } else {
      double x = t0, h;
      // ---- k1 = f(t0, y0) (also f0 of hinit); the barrier inside orders the LDS fills above
      LB_RHS(y_r, y_i, k_r[1], k_i[1], L.coef);
      if (isfinite(C.dt) && C.dt > 0.0) {
        h = fmin(C.dt, t1 - t0);
      } else {
        // hinit (rust/src/ode/common.rs:72-118)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double sr = C.abstol + fabs(y_r[r]) * C.reltol, si = C.abstol + fabs(y_i[r]) * C.reltol;
          const double a0 = y_r[r] / sr, a1 = y_i[r] / si, b0 = k_r[1][r] / sr, b1 = k_i[1][r] / si;
          const bool ok = S.valid(r), cpx_ = !S.diag(r);
          block_store(P, tid, W, r, L.red, 0, ok ? a0 * a0 + (cpx_ ? a1 * a1 : 0.0) : 0.0);
          block_store(P, tid, W, r, L.red, 1, ok ? b0 * b0 + (cpx_ ? b1 * b1 : 0.0) : 0.0);
        }
        __syncthreads();
        const double d0 = block_total(P, L.red, 0), d1 = block_total(P, L.red, 1), hmax = fabs(t1 - x);
        double h0 = (d0 < 1.0e-10 || d1 < 1.0e-10) ? 1.0e-6 : 0.01 * sqrt(d0 / d1);
        h0 = fmin(h0, hmax);
        if (tid == 0) eval_coefficients(P, L.slots + P.n_slots, L.coef + KK, x + h0, true, L.vm);
        if constexpr (G > 0) __syncthreads();
        double yt_r[R], yt_i[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          yt_r[r] = y_r[r] + k_r[1][r] * h0;
          yt_i[r] = y_i[r] + k_i[1][r] * h0;
        }
        LB_RHS(yt_r, yt_i, k_r[2], k_i[2], L.coef + KK);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double sr = C.abstol + fabs(y_r[r]) * C.reltol, si = C.abstol + fabs(y_i[r]) * C.reltol;
          const double a0 = (k_r[2][r] - k_r[1][r]) / sr, a1 = (k_i[2][r] - k_i[1][r]) / si;
          block_store(P, tid, W, r, L.red, 2, S.valid(r) ? a0 * a0 + (S.diag(r) ? 0.0 : a1 * a1) : 0.0);
        }
        __syncthreads();
        const double d2 = sqrt(block_total(P, L.red, 2)) / h0;
        const double h1 = fmax(sqrt(d1), fabs(d2)) <= 1.0e-15 ? fmax(1.0e-6, fabs(h0) * 1.0e-3)
                                                              : pow(0.01 / fmax(sqrt(d1), d2), 1.0 / 5.0);
        h = fmin(100.0 * fabs(h0), fmin(h1, hmax));
      }
      rhs_calls = 1;
      if constexpr (EV) {
        if (C.event_kind) {
          bool cpx = false;
          ev_old = event_value<R>(P, C, L, S, tid, W, t0, L.slots, y_r, cpx);
          if (cpx) { status = LB_TRAJ_EVENT_COMPLEX; fail_x = t0; }
        }
      }
      // ---- points at t0 (dopri5.rs:212-220)
      if (!step_push) {
        while (si < C.n_saveat && fabs(C.saveat[si] - t0) <= time_tol(t0)) {
          push(C.saveat[si], y_r, y_i, true);
          ++si;
        }
      } else if (C.save_start) {
        push(t0, y_r, y_i, true);
      }
      // ---- controller state (common.rs:32-70)
      const double beta = 0.04, alpha = 0.2 - beta * 0.75, facc1 = 1.0 / 0.2, facc2 = 1.0 / 10.0;
      const double safety = 0.9, hmax = fabs(t1 - t0);
      double fac_old = 1.0e-4;
      bool reject = false, last = false;
      int64_t ns = 0, nst = 0, ia = 0;

      while (!last && (!EV || status == 0)) {
        if (ns > C.maxiters) { status = LB_TRAJ_MAXITERS; fail_x = x; break; }
        bool clipped = false;
        if (C.replay_h) {
          // recorded schedule: step sizes of a tape, every step accepted
          const int64_t nrep = C.replay_n[traj];
          if (accepted >= nrep || accepted >= C.tape_cap) { status = LB_TRAJ_CAPACITY; fail_x = x; break; }
          h = C.replay_h[traj * C.tape_cap + accepted];
          last = accepted + 1 == nrep;
        } else {
          if (0.1 * fabs(h) <= DBL_EPSILON * fabs(x)) { status = LB_TRAJ_UNDERFLOW; fail_x = x; break; }
          if (x + 1.01 * h - t1 > 0.0) { h = t1 - x; last = true; clipped = true; }
        }
        ++ns;
        // The expression machine needs registers of its own: the single-wave
        // form parks its state in LDS meanwhile (the multi-wave form keeps
        // it in the other waves' registers).
        if constexpr (G > 0) {
          L.park[tid] = double2{y_r[0], y_i[0]};
          L.park[64 + tid] = double2{k_r[1][0], k_i[1][0]};
        }
        // time-dependent coefficients at the six stage times, one lane per stage
        if (P.time_dependent && tid < 6) {
          const int s = tid + 2;
          eval_coefficients(P, L.slots + s * P.n_slots, L.coef + s * KK, x + h * T::c(s), true,
                            L.vm + tid * P.vm_depth);
        }
        if constexpr (G > 0) {
          __syncthreads();  // stage coefficients are read by every lane
          const double2 py = L.park[tid], pk = L.park[64 + tid];
          y_r[0] = py.x;
          y_i[0] = py.y;
          k_r[1][0] = pk.x;
          k_i[1][0] = pk.y;
        }
        // the stage vectors of the previous step are dead from here on
#pragma unroll
        for (int j = 2; j <= 7; ++j)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            k_r[j][r] = 0.0;
            k_i[j][r] = 0.0;
          }
        const bool need_stiff = ((accepted + 1) % 1000 == 0) || ia > 0;
        double yt_r[R], yt_i[R], yn_r[R], yn_i[R];
        // Stages 2..7 share one RHS site (a rolled loop keeps the step's code
        // small enough for the instruction cache); stage 7 is the solution,
        // whose RHS is the FSAL derivative.
#pragma unroll 1
        for (int s = 2; s <= 7; ++s) {
          switch (s) {
            case 2: stage_combo<R, SOLVER, 2>(h, y_r, y_i, k_r, k_i, yt_r, yt_i); break;
            case 3: stage_combo<R, SOLVER, 3>(h, y_r, y_i, k_r, k_i, yt_r, yt_i); break;
            case 4: stage_combo<R, SOLVER, 4>(h, y_r, y_i, k_r, k_i, yt_r, yt_i); break;
            case 5: stage_combo<R, SOLVER, 5>(h, y_r, y_i, k_r, k_i, yt_r, yt_i); break;
            case 6: stage_combo<R, SOLVER, 6>(h, y_r, y_i, k_r, k_i, yt_r, yt_i); break;
            default:
              stage_combo<R, SOLVER, 7>(h, y_r, y_i, k_r, k_i, yn_r, yn_i);
              if (need_stiff) {
                // |yn - y_stage6|^2 per block (dopri5.rs:255-264)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                  const double a0 = yn_r[r] - yt_r[r], a1 = yn_i[r] - yt_i[r];
                  block_store(P, tid, W, r, L.red, 3, a0 * a0 + a1 * a1);
                }
              }
#pragma unroll
              for (int r = 0; r < R; ++r) {
                yt_r[r] = yn_r[r];
                yt_i[r] = yn_i[r];
              }
              break;
          }
          double ko_r[R], ko_i[R];
          LB_RHS(yt_r, yt_i, ko_r, ko_i, L.coef + s * KK);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            switch (s) {
              case 2: k_r[2][r] = ko_r[r]; k_i[2][r] = ko_i[r]; break;
              case 3: k_r[3][r] = ko_r[r]; k_i[3][r] = ko_i[r]; break;
              case 4: k_r[4][r] = ko_r[r]; k_i[4][r] = ko_i[r]; break;
              case 5: k_r[5][r] = ko_r[r]; k_i[5][r] = ko_i[r]; break;
              case 6: k_r[6][r] = ko_r[r]; k_i[6][r] = ko_i[r]; break;
              default: k_r[7][r] = ko_r[r]; k_i[7][r] = ko_i[r]; break;
            }
          }
        }
        rhs_calls += 6;
        // ---- error norm (dopri5.rs:88-102, tsit5.rs:104-119), integrand, stiffness
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double er = T::e(1) * k_r[1][r], ei = T::e(1) * k_i[1][r];
#pragma unroll
          for (int j = 2; j <= 7; ++j) {
            if (T::e(j) != 0.0) {
              er = fma(T::e(j), k_r[j][r], er);
              ei = fma(T::e(j), k_i[j][r], ei);
            }
          }
          er *= h;
          ei *= h;
          const double sr = C.abstol + fmax(fabs(y_r[r]), fabs(yn_r[r])) * C.reltol;
          const double si = C.abstol + fmax(fabs(y_i[r]), fabs(yn_i[r])) * C.reltol;
          const double qr = er / sr, qi = ei / si;
          block_store(P, tid, W, r, L.red, 0, S.valid(r) ? qr * qr + (S.diag(r) ? 0.0 : qi * qi) : 0.0);
          if (C.want_integral && step_push)
            block_store(P, tid, W, r, L.red, 1, S.diag(r) && S.valid(r) ? L.wts[S.state(r)] * yn_r[r] : 0.0);
          if (need_stiff) {
            const double a0 = k_r[7][r] - k_r[6][r], a1 = k_i[7][r] - k_i[6][r];
            block_store(P, tid, W, r, L.red, 2, a0 * a0 + a1 * a1);
          }
        }
        __syncthreads();
        const double err = C.replay_h ? 0.0 : sqrt(block_total(P, L.red, 0) / (double)P.dim);
        // ---- controller (common.rs:48-69); the two powers at once in each
        // wave (lanes 0-31: err^alpha, lanes 32-63: fac_old^-beta)
        const int wl = tid & 63;
        const double pw = pow(wl < 32 ? err : fac_old, wl < 32 ? alpha : -beta);
        const double fac11 = __shfl(pw, 0, 64), fo = __shfl(pw, 32, 64);
        double fac = fac11 * fo;
        const double fac_s = fac / safety;
        fac = fmax(facc2, fmin(facc1, fac_s));
        double hn = h / fac;
        if (C.tape_h && !C.replay_h) {
          // attempt tape: what the adjoint needs to differentiate this attempt
          // and the controller's decision (one record per attempt)
          const int64_t a = ns - 1;
          if (a < C.tape_cap) {
            if (tid == 0) {
              int fl = clipped ? LB_TAPE_CLIPPED : 0;
              double h_next;
              if (err <= 1.0) {
                h_next = hn;
                fl |= LB_TAPE_ACCEPTED | (accepted == 0 ? LB_TAPE_FIRST : 0) | (err > 1.0e-4 ? LB_TAPE_ERRQ : 0);
                if (!(fac_s > facc2 && fac_s < facc1)) fl |= LB_TAPE_CLAMPED;
                double hc = hn;
                if (fabs(hc) > hmax) { hc = hmax; fl |= LB_TAPE_HMAX; }
                if (reject && fabs(h) < fabs(hc)) fl |= LB_TAPE_MINRULE;
              } else {
                const double fr = fac11 / safety;
                h_next = h / fmin(facc1, fr);
                if (!(fr < facc1)) fl |= LB_TAPE_CLAMPED;
              }
              const int64_t o = traj * C.tape_cap + a;
              C.tape_t[o] = x;
              C.tape_h[o] = h;
              C.tape_c[3 * o] = err;
              C.tape_c[3 * o + 1] = h_next;
              C.tape_c[3 * o + 2] = fac_old;
              C.tape_f[o] = fl;
            }
            if (a % C.ckpt_every == 0) {
              const int64_t ncap = (C.tape_cap + C.ckpt_every - 1) / C.ckpt_every;
              double* ck = C.tape_y + ((traj * ncap + a / C.ckpt_every) * P.S) * 2;
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int s = tid + W * r;
                if (s < P.S) {
                  ck[2 * s] = y_r[r];
                  ck[2 * s + 1] = y_i[r];
                }
              }
            }
          }
        }
        if (err <= 1.0) {
          fac_old = fmax(err, 1.0e-4);
          if (fabs(hn) > hmax) hn = hmax;
          if (reject) hn = fmin(fabs(hn), fabs(h));
          reject = false;
          ++accepted;
          if (accepted % 1000 == 0 || ia > 0) {
            const double n2 = block_total(P, L.red, 2), dd2 = block_total(P, L.red, 3);
            const double hl = dd2 > 0.0 ? h * sqrt(n2 / dd2) : 0.0;
            if (hl > 3.25) {
              ++ia;
              nst = 0;
              if (ia == 15) { status = LB_TRAJ_STIFF; fail_x = x; break; }
            } else {
              ++nst;
              if (nst == 6) ia = 0;
            }
          }
          const double xo = x;
          x += h;
          // ---- terminal stop event (dopri5.rs:280-317, tsit5.rs:274-309): g at
          // the step end (the slots of the last stage are those at x), a root
          // by 60 bisections on the dense output, its state left in yt
          bool ev_hit = false;
          double te = x;
          if constexpr (EV) {
            if (C.event_kind) {
              bool cpx = false;
              const double g1 = event_value<R>(P, C, L, S, tid, W, x, L.slots + 7 * P.n_slots, yn_r, cpx);
              if (cpx) { status = LB_TRAJ_EVENT_COMPLEX; fail_x = x; break; }
              if (crossed_zero(ev_old, g1)) {
                ev_hit = true;
                if (ev_old == 0.0 || g1 == 0.0) {
                  te = ev_old == 0.0 ? xo : x;
#pragma unroll
                  for (int r = 0; r < R; ++r) {
                    yt_r[r] = ev_old == 0.0 ? y_r[r] : yn_r[r];
                    yt_i[r] = ev_old == 0.0 ? y_i[r] : yn_i[r];
                  }
                } else {
                  double lo = 0.0, hi = 1.0, glo = ev_old;
                  for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    dense_point<R, SOLVER>(mid, h, y_r, y_i, yn_r, yn_i, k_r, k_i, yt_r, yt_i);
                    // slot copy 0 is free after the trajectory's start
                    if (C.event_kind == LB_EVENT_EXPRESSION && tid == 0) eval_compounds(P, L.slots, xo + mid * h, L.vm);
                    const double gm = event_value<R>(P, C, L, S, tid, W, xo + mid * h, L.slots, yt_r, cpx);
                    if (cpx) { fail_x = xo + mid * h; break; }
                    if (gm == 0.0) { lo = mid; hi = mid; break; }
                    if ((glo < 0.0 && gm > 0.0) || (glo > 0.0 && gm < 0.0)) {
                      hi = mid;
                    } else {
                      lo = mid;
                      glo = gm;
                    }
                  }
                  if (cpx) { status = LB_TRAJ_EVENT_COMPLEX; break; }
                  const double th = 0.5 * (lo + hi);
                  dense_point<R, SOLVER>(th, h, y_r, y_i, yn_r, yn_i, k_r, k_i, yt_r, yt_i);
                  te = xo + th * h;
                }
              }
              ev_old = g1;
            }
          }
          if (step_push) {
            if (!ev_hit) {
              if (C.want_integral) {
                const double v = block_total(P, L.red, 1);
                if (isfinite(last_t)) integral += 0.5 * (last_v + v) * (x - last_t);
                last_t = x;
                last_v = v;
              }
              if constexpr (EV) last_pt = x;
              if (C.trace) push(x, yn_r, yn_i, rate);
            }
          } else {
            // save plan points inside (xo, x], or up to the root (dopri5.rs:318-344)
            const double tb = ev_hit ? te : x;
            while (si < C.n_saveat && C.saveat[si] <= tb + time_tol(tb)) {
              const double ts = C.saveat[si];
              if (ts >= xo - time_tol(xo)) {
                if (fabs(ts - xo) <= time_tol(xo)) {
                  push(ts, y_r, y_i, true);
                } else if (fabs(ts - x) <= time_tol(x)) {
                  push(ts, yn_r, yn_i, true);
                } else {
                  double ds_r[R], ds_i[R];
                  dense_point<R, SOLVER>((ts - xo) / h, h, y_r, y_i, yn_r, yn_i, k_r, k_i, ds_r, ds_i);
                  push(ts, ds_r, ds_i, true);
                }
              }
              ++si;
            }
          }
          if (EV && ev_hit) {
            if (!(isfinite(last_pt) && fabs(last_pt - te) <= time_tol(te))) push(te, yt_r, yt_i, true);
            // A final state output is the root's state; the reference's save
            // plan for it is {t1}, so a root within time_tol of t1 leaves the
            // state pushed at t1 there instead (dopri5.rs:318-352).
            if (!C.trace && C.output_kind != LB_OUT_INTEGRAL && step_push && t1 <= te + time_tol(te)) {
              if (fabs(t1 - xo) <= time_tol(xo)) {
#pragma unroll
                for (int r = 0; r < R; ++r) { yt_r[r] = y_r[r]; yt_i[r] = y_i[r]; }
              } else if (fabs(t1 - x) <= time_tol(x)) {
#pragma unroll
                for (int r = 0; r < R; ++r) { yt_r[r] = yn_r[r]; yt_i[r] = yn_i[r]; }
              } else {
                dense_point<R, SOLVER>((t1 - xo) / h, h, y_r, y_i, yn_r, yn_i, k_r, k_i, yt_r, yt_i);
              }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
              y_r[r] = yt_r[r];
              y_i[r] = yt_i[r];
            }
            ev_fired = true;
            ev_t = te;
            break;
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            y_r[r] = yn_r[r];
            y_i[r] = yn_i[r];
            k_r[1][r] = k_r[7][r];
            k_i[1][r] = k_i[7][r];
          }
          if (last) break;
        } else {
          hn = h / fmin(facc1, fac11 / safety);
          reject = true;
          last = false;
          if (accepted >= 1) ++rejected;
        }
        h = hn;
      }
      attempts = ns;
    }
#undef LB_RHS
#undef PUB

    // ---- outputs
    if constexpr (EV) {
      if (C.out_event_t && tid == 0) C.out_event_t[traj] = ev_t;
    }
    if (status == 0 && !step_push && si != C.n_saveat && (!EV || !ev_fired)) {
      status = LB_TRAJ_SAVEAT;
      fail_x = (double)si;  // plan points reached
    }
    if (status == 0 && C.trace && n_pushed > C.capacity) status = LB_TRAJ_CAPACITY;
    if (status == 0 && C.tape_h && !C.replay_h && attempts > C.tape_cap) status = LB_TRAJ_CAPACITY;
    if (C.tape_n && tid == 0) C.tape_n[traj] = (int32_t)attempts;
    const bool ok = status == 0;
    const double nanv = __builtin_nan("");
    if (C.trace) {
      // rows were written as the points were pushed; dropped elements stay
      // at the zeros the caller provides
      if (tid == 0) {
        C.out_counts[traj] = (int32_t)n_pushed;
        if (C.out_integral) C.out_integral[traj] = ok ? integral : nanv;
        C.out_stats[4 * traj + 0] = accepted;
        C.out_stats[4 * traj + 1] = rejected;
        C.out_stats[4 * traj + 2] = rhs_calls;
        C.out_stats[4 * traj + 3] = status;
        C.out_fail_t[traj] = fail_x;
      }
      continue;
    }
    if (P.restricted) {
      // elements outside the restricted plan are exactly zero
      if (C.output_kind == LB_OUT_POPULATIONS) {
        for (int i = tid; i < P.n; i += W)
          if (!P.upper_active[P.packed_upper[i]]) C.out_values[traj * P.n + i] = ok ? 0.0 : nanv;
      } else if (C.output_kind == LB_OUT_FULL) {
        for (int k = tid; k < P.dim; k += W)
          if (!P.upper_active[P.packed_upper[k]]) C.out_values[traj * P.dim + k] = ok ? 0.0 : nanv;
      } else if (C.output_kind == LB_OUT_SELECTED) {
        for (int q = tid; q < C.n_selected; q += W)
          if (!P.upper_active[C.sel_upper[q]]) {
            C.out_values[2 * (traj * C.n_selected + q)] = ok ? 0.0 : nanv;
            C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? 0.0 : nanv;
          }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!S.valid(r)) continue;
      const int s = tid + W * r;
      if (C.output_kind == LB_OUT_POPULATIONS) {
        if (S.diag(r)) C.out_values[traj * P.n + S.state(r)] = ok ? y_r[r] : nanv;
      } else if (C.output_kind == LB_OUT_FULL) {
        C.out_values[traj * P.dim + P.slot_re[s]] = ok ? y_r[r] : nanv;
        if (!S.diag(r)) C.out_values[traj * P.dim + P.slot_im[s]] = ok ? y_i[r] : nanv;
      } else if (C.output_kind == LB_OUT_SELECTED) {
        const int up = P.slot_up[s];
        for (int q = 0; q < C.n_selected; ++q) {
          if (C.sel_upper[q] != up) continue;
          const double im = (C.sel_flags[q] & 1) ? 0.0 : ((C.sel_flags[q] & 2) ? -y_i[r] : y_i[r]);
          C.out_values[2 * (traj * C.n_selected + q)] = ok ? y_r[r] : nanv;
          C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? im : nanv;
        }
      }
    }
    if (tid == 0) {
      if (C.output_kind == LB_OUT_INTEGRAL) C.out_values[traj] = ok ? integral : nanv;
      if (C.out_integral) C.out_integral[traj] = ok ? integral : nanv;
      C.out_stats[4 * traj + 0] = accepted;
      C.out_stats[4 * traj + 1] = rejected;
      C.out_stats[4 * traj + 2] = rhs_calls;
      C.out_stats[4 * traj + 3] = status;
      C.out_fail_t[traj] = fail_x;
    }
  }
}

// ---------------------------------------------------------------------------
// Fixed-step integrators (rust/src/ode/fixed.rs:133-249): steps of dt,
// shortened to land on t1 and on the next save-plan point, Dopri5 (6 RHS),
// midpoint RK2 (2) or classical RK4 (4) stages with the stage sums written in
// the reference's operation order, no error control; a push after every step
// or at the save-plan points.  Persistent multi-wave form (any plan size).
// ---------------------------------------------------------------------------
namespace {

template <int ST>  // 0 Dopri5, 1 RK2, 2 RK4
struct FixedStep {
  static constexpr int NS = ST == 0 ? 6 : ST == 1 ? 2 : 4;
  // time of stage s (1-based), as the reference writes it
  __device__ static double time(double x, double h, int s) {
    if constexpr (ST == 1) return s == 1 ? x : x + 0.5 * h;
    else if constexpr (ST == 2) return s == 1 ? x : s == 4 ? x + h : x + 0.5 * h;
    else
      return s == 1 ? x : s == 2 ? x + h * (1.0 / 5.0) : s == 3 ? x + h * (3.0 / 10.0)
           : s == 4 ? x + h * (4.0 / 5.0) : s == 5 ? x + h * (8.0 / 9.0) : x + h;
  }
};

// Outputs of a finished trajectory (final values, integral, stats), as the
// adaptive kernel writes them.
template <int R>
__device__ void write_final(const DevPlan& P, const DevCall& C, const Slots<R>& S, int tid, int W, int64_t traj,
                            const double (&y_r)[R], const double (&y_i)[R], int status, double fail_x,
                            double integral, int64_t accepted, int64_t rhs_calls, int64_t n_pushed) {
  const bool ok = status == 0;
  const double nanv = __builtin_nan("");
  if (C.trace) {
    if (tid == 0) {
      C.out_counts[traj] = (int32_t)n_pushed;
      if (C.out_integral) C.out_integral[traj] = ok ? integral : nanv;
      C.out_stats[4 * traj + 0] = accepted;
      C.out_stats[4 * traj + 1] = 0;
      C.out_stats[4 * traj + 2] = rhs_calls;
      C.out_stats[4 * traj + 3] = status;
      C.out_fail_t[traj] = fail_x;
    }
    return;
  }
  if (P.restricted) {
    if (C.output_kind == LB_OUT_POPULATIONS) {
      for (int i = tid; i < P.n; i += W)
        if (!P.upper_active[P.packed_upper[i]]) C.out_values[traj * P.n + i] = ok ? 0.0 : nanv;
    } else if (C.output_kind == LB_OUT_FULL) {
      for (int k = tid; k < P.dim; k += W)
        if (!P.upper_active[P.packed_upper[k]]) C.out_values[traj * P.dim + k] = ok ? 0.0 : nanv;
    } else if (C.output_kind == LB_OUT_SELECTED) {
      for (int q = tid; q < C.n_selected; q += W)
        if (!P.upper_active[C.sel_upper[q]]) {
          C.out_values[2 * (traj * C.n_selected + q)] = ok ? 0.0 : nanv;
          C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? 0.0 : nanv;
        }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!S.valid(r)) continue;
    const int s = tid + W * r;
    if (C.output_kind == LB_OUT_POPULATIONS) {
      if (S.diag(r)) C.out_values[traj * P.n + S.state(r)] = ok ? y_r[r] : nanv;
    } else if (C.output_kind == LB_OUT_FULL) {
      C.out_values[traj * P.dim + P.slot_re[s]] = ok ? y_r[r] : nanv;
      if (!S.diag(r)) C.out_values[traj * P.dim + P.slot_im[s]] = ok ? y_i[r] : nanv;
    } else if (C.output_kind == LB_OUT_SELECTED) {
      const int up = P.slot_up[s];
      for (int q = 0; q < C.n_selected; ++q) {
        if (C.sel_upper[q] != up) continue;
        const double im = (C.sel_flags[q] & 1) ? 0.0 : ((C.sel_flags[q] & 2) ? -y_i[r] : y_i[r]);
        C.out_values[2 * (traj * C.n_selected + q)] = ok ? y_r[r] : nanv;
        C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? im : nanv;
      }
    }
  }
  if (tid == 0) {
    if (C.output_kind == LB_OUT_INTEGRAL) C.out_values[traj] = ok ? integral : nanv;
    if (C.out_integral) C.out_integral[traj] = ok ? integral : nanv;
    C.out_stats[4 * traj + 0] = accepted;
    C.out_stats[4 * traj + 1] = 0;
    C.out_stats[4 * traj + 2] = rhs_calls;
    C.out_stats[4 * traj + 3] = status;
    C.out_fail_t[traj] = fail_x;
  }
}

}  // namespace

template <int R, int ST, int MAXW, int WAVES, bool EV = false>
__global__ __launch_bounds__(MAXW, WAVES) void lb_fixed_kernel(const DevPlan P, const DevCall C) {
  using F = FixedStep<ST>;
  constexpr int NS = F::NS;
  const int tid = threadIdx.x, W = blockDim.x;
  const Lds L = carve(P);
  const int KK = P.K + 1;
  Slots<R> S;
  load_slots<R>(P, S, tid, W);
  stage_tables(P, L, tid, W);

  for (;;) {
    __syncthreads();
    if (tid == 0) L.misc[0] = atomicAdd(C.queue, 1);
    __syncthreads();
    const int64_t local = L.misc[0];
    if (local >= C.n_traj) break;
    const int64_t traj = C.queue_order ? (int64_t)C.queue_order[local + C.traj_offset] : local + C.traj_offset;
    if (traj < 0 || traj >= C.n_total) continue;
    const double t0 = C.t0;
    const double t1 = C.t1_per_traj ? C.t1_per_traj[traj] : C.t1;

    // ---- parameters and coefficients at t0, self-term folds, weights
    init_slots(P, L.slots, tid, W);
    if (tid == 0) {
      apply_overrides(C, traj, L.slots);
      eval_coefficients(P, L.slots, L.coef, t0, false, L.vm);
    }
    for (int i = tid; i < P.n; i += W) {
      double w = 0.0;
      for (int q = 0; q < C.n_weights; ++q)
        if (C.weight_index[q] == i) w += C.weight_value[q];
      L.wts[i] = w;
      if constexpr (EV) L.evv[i] = 0.0;
    }
    __syncthreads();
    for (int i = tid; i < 7 * P.n_slots; i += W) L.slots[P.n_slots + i] = L.slots[i % P.n_slots];
    for (int i = tid; i < 7 * KK; i += W) L.coef[KK + i] = L.coef[i % KK];
    fold_self(P, tid, W, L.coef, L.selfc);

    double y_r[R], y_i[R];
    const double* y0 = C.y0 + (C.y0_shared ? 0 : traj * (int64_t)P.dim);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int s = tid + W * r;
      y_r[r] = S.valid(r) ? y0[P.slot_re[s]] : 0.0;
      y_i[r] = (S.valid(r) && !S.diag(r)) ? y0[P.slot_im[s]] : 0.0;
    }
    bool bad_support = false;
    if (P.restricted) {
      if (tid == 0) L.misc[1] = 0;
      __syncthreads();
      for (int k = tid; k < P.dim; k += W)
        if (!P.upper_active[P.packed_upper[k]] && y0[k] != 0.0) L.misc[1] = 1;
      __syncthreads();
      bad_support = L.misc[1] != 0;
    }

    int64_t accepted = 0, rhs_calls = 0, n_pushed = 0, si = 0;
    int status = 0;
    double fail_x = 0.0, integral = 0.0, last_t = __builtin_nan(""), last_v = 0.0;
    double ev_old = 0.0, ev_t = __builtin_nan("");  // stop event: g at the last step end, root time
    bool ev_fired = false;
    const bool step_push = C.n_saveat == 0;
    const bool rate = C.output_kind == LB_OUT_RATE;
    auto push = [&](double tp, const double (&vr)[R], const double (&vi)[R]) {
      double v = last_v;
      if (C.want_integral || rate) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          block_store(P, tid, W, r, L.red, 1, S.diag(r) && S.valid(r) ? L.wts[S.state(r)] * vr[r] : 0.0);
        __syncthreads();
        v = block_total(P, L.red, 1);
        __syncthreads();
        if (C.want_integral) {
          if (isfinite(last_t)) integral += 0.5 * (last_v + v) * (tp - last_t);
          last_t = tp;
          last_v = v;
        }
      }
      if (C.trace) {
        if (n_pushed < C.capacity) {
          const int64_t row = traj * C.capacity + n_pushed;
          if (C.output_kind == LB_OUT_INTEGRAL) {
            if (tid == 0) C.out_values[row] = integral;
          } else if (rate) {
            if (tid == 0) C.out_values[row] = v;
          } else {
            write_row<R>(P, C, S, tid, W, row, vr, vi);
          }
          if (tid == 0 && C.out_times) C.out_times[row] = tp;
        }
        ++n_pushed;
      }
    };
    double k_r[7][R], k_i[7][R];
    int buf = 0;
#define LB_FRHS(xr, xi, kr, ki, cf)                                         \
  do {                                                                      \
    rhs_eval<R>(P, L, S, xr, xi, kr, ki, L.pub + buf * 2 * P.n_pub, cf, tid, W); \
    buf ^= 1;                                                               \
  } while (0)

    if (bad_support) {
      status = LB_TRAJ_BAD_SUPPORT;
      fail_x = t0;
    } else if (t1 < t0) {
      status = LB_TRAJ_BAD_SPAN;
      fail_x = t0;
    } else if (!(isfinite(C.dt) && C.dt > 0.0)) {
      status = LB_TRAJ_BAD_DT;
      fail_x = t0;
    } else if (t1 == t0) {
      if (C.save_start) push(t0, y_r, y_i);
    } else {
      if (!step_push) {
        while (si < C.n_saveat && fabs(C.saveat[si] - t0) <= time_tol(t0)) {
          push(C.saveat[si], y_r, y_i);
          ++si;
        }
      } else if (C.save_start) {
        push(t0, y_r, y_i);
      }
      if constexpr (EV) {
        bool cpx = false;
        ev_old = event_value<R>(P, C, L, S, tid, W, t0, L.slots, y_r, cpx);
        if (cpx) { status = LB_TRAJ_EVENT_COMPLEX; fail_x = t0; }
      }
      double x = t0;
      int64_t steps = 0;
      while ((!EV || status == 0) && x < t1 - time_tol(t1)) {
        if (steps >= C.maxiters) { status = LB_TRAJ_MAXITERS; fail_x = x; break; }
        double target = fmin(x + C.dt, t1);
        if (!step_push && si < C.n_saveat && C.saveat[si] > x + time_tol(x)) target = fmin(target, C.saveat[si]);
        const double h = target - x;
        if (h <= 0.0 || 0.1 * h <= DBL_EPSILON * fmax(fabs(x), 1.0)) { status = LB_TRAJ_UNDERFLOW; fail_x = x; break; }
        __syncthreads();  // the previous step's RHS have read their coefficient rows
        if (P.time_dependent && tid < NS)
          eval_coefficients(P, L.slots + (tid + 1) * P.n_slots, L.coef + (tid + 1) * KK, F::time(x, h, tid + 1),
                            true, L.vm + tid * P.vm_depth);
        // (the barrier inside the first RHS orders the rows before the gathers)
        double yt_r[R], yt_i[R], yn_r[R], yn_i[R];
        LB_FRHS(y_r, y_i, k_r[1], k_i[1], L.coef + KK);
        if constexpr (ST == 1) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yt_r[r] = y_r[r] + 0.5 * h * k_r[1][r];
            yt_i[r] = y_i[r] + 0.5 * h * k_i[1][r];
          }
          LB_FRHS(yt_r, yt_i, k_r[2], k_i[2], L.coef + 2 * KK);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yn_r[r] = y_r[r] + h * k_r[2][r];
            yn_i[r] = y_i[r] + h * k_i[2][r];
          }
        } else if constexpr (ST == 2) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yt_r[r] = y_r[r] + 0.5 * h * k_r[1][r];
            yt_i[r] = y_i[r] + 0.5 * h * k_i[1][r];
          }
          LB_FRHS(yt_r, yt_i, k_r[2], k_i[2], L.coef + 2 * KK);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yt_r[r] = y_r[r] + 0.5 * h * k_r[2][r];
            yt_i[r] = y_i[r] + 0.5 * h * k_i[2][r];
          }
          LB_FRHS(yt_r, yt_i, k_r[3], k_i[3], L.coef + 3 * KK);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yt_r[r] = y_r[r] + h * k_r[3][r];
            yt_i[r] = y_i[r] + h * k_i[3][r];
          }
          LB_FRHS(yt_r, yt_i, k_r[4], k_i[4], L.coef + 4 * KK);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yn_r[r] = y_r[r] + h * (k_r[1][r] + 2.0 * k_r[2][r] + 2.0 * k_r[3][r] + k_r[4][r]) / 6.0;
            yn_i[r] = y_i[r] + h * (k_i[1][r] + 2.0 * k_i[2][r] + 2.0 * k_i[3][r] + k_i[4][r]) / 6.0;
          }
        } else {
#define LB_FSUM(expr_r, expr_i)            \
  _Pragma("unroll") for (int r = 0; r < R; ++r) { \
    yt_r[r] = y_r[r] + h * (expr_r);       \
    yt_i[r] = y_i[r] + h * (expr_i);       \
  }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yt_r[r] = y_r[r] + h * (1.0 / 5.0) * k_r[1][r];
            yt_i[r] = y_i[r] + h * (1.0 / 5.0) * k_i[1][r];
          }
          LB_FRHS(yt_r, yt_i, k_r[2], k_i[2], L.coef + 2 * KK);
          LB_FSUM((3.0 / 40.0) * k_r[1][r] + (9.0 / 40.0) * k_r[2][r],
                  (3.0 / 40.0) * k_i[1][r] + (9.0 / 40.0) * k_i[2][r])
          LB_FRHS(yt_r, yt_i, k_r[3], k_i[3], L.coef + 3 * KK);
          LB_FSUM((44.0 / 45.0) * k_r[1][r] - (56.0 / 15.0) * k_r[2][r] + (32.0 / 9.0) * k_r[3][r],
                  (44.0 / 45.0) * k_i[1][r] - (56.0 / 15.0) * k_i[2][r] + (32.0 / 9.0) * k_i[3][r])
          LB_FRHS(yt_r, yt_i, k_r[4], k_i[4], L.coef + 4 * KK);
          LB_FSUM((19372.0 / 6561.0) * k_r[1][r] - (25360.0 / 2187.0) * k_r[2][r] + (64448.0 / 6561.0) * k_r[3][r] -
                      (212.0 / 729.0) * k_r[4][r],
                  (19372.0 / 6561.0) * k_i[1][r] - (25360.0 / 2187.0) * k_i[2][r] + (64448.0 / 6561.0) * k_i[3][r] -
                      (212.0 / 729.0) * k_i[4][r])
          LB_FRHS(yt_r, yt_i, k_r[5], k_i[5], L.coef + 5 * KK);
          LB_FSUM((9017.0 / 3168.0) * k_r[1][r] - (355.0 / 33.0) * k_r[2][r] + (46732.0 / 5247.0) * k_r[3][r] +
                      (49.0 / 176.0) * k_r[4][r] - (5103.0 / 18656.0) * k_r[5][r],
                  (9017.0 / 3168.0) * k_i[1][r] - (355.0 / 33.0) * k_i[2][r] + (46732.0 / 5247.0) * k_i[3][r] +
                      (49.0 / 176.0) * k_i[4][r] - (5103.0 / 18656.0) * k_i[5][r])
          LB_FRHS(yt_r, yt_i, k_r[6], k_i[6], L.coef + 6 * KK);
#undef LB_FSUM
#pragma unroll
          for (int r = 0; r < R; ++r) {
            yn_r[r] = y_r[r] + h * ((35.0 / 384.0) * k_r[1][r] + (500.0 / 1113.0) * k_r[3][r] +
                                    (125.0 / 192.0) * k_r[4][r] - (2187.0 / 6784.0) * k_r[5][r] +
                                    (11.0 / 84.0) * k_r[6][r]);
            yn_i[r] = y_i[r] + h * ((35.0 / 384.0) * k_i[1][r] + (500.0 / 1113.0) * k_i[3][r] +
                                    (125.0 / 192.0) * k_i[4][r] - (2187.0 / 6784.0) * k_i[5][r] +
                                    (11.0 / 84.0) * k_i[6][r]);
          }
        }
        rhs_calls += NS;
        ++accepted;
        ++steps;
        const double xo = x;
        x += h;
        // terminal stop event (fixed.rs:213-238): the root is the step's end
        // (its start when g was zero there), with the end state
        bool ev_hit = false;
        if constexpr (EV) {
          // slot copy 0 is free after the trajectory's start
          if (C.event_kind == LB_EVENT_EXPRESSION && tid == 0) eval_compounds(P, L.slots, x, L.vm);
          bool cpx = false;
          const double g1 = event_value<R>(P, C, L, S, tid, W, x, L.slots, yn_r, cpx);
          if (cpx) { status = LB_TRAJ_EVENT_COMPLEX; fail_x = x; break; }
          if (crossed_zero(ev_old, g1)) {
            ev_hit = true;
            ev_t = ev_old == 0.0 ? xo : x;
          }
          ev_old = g1;
        }
        if (step_push) {
          if (!ev_hit) push(x, yn_r, yn_i);
        } else {
          const double tb = ev_hit ? ev_t : x;
          while (si < C.n_saveat && C.saveat[si] <= tb + time_tol(tb)) {
            push(C.saveat[si], yn_r, yn_i);
            ++si;
          }
        }
        if (EV && ev_hit) {
          push(ev_t, yn_r, yn_i);
          ev_fired = true;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            y_r[r] = yn_r[r];
            y_i[r] = yn_i[r];
          }
          break;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          y_r[r] = yn_r[r];
          y_i[r] = yn_i[r];
        }
      }
      if (status == 0 && !step_push && si != C.n_saveat && (!EV || !ev_fired)) {
        status = LB_TRAJ_SAVEAT;
        fail_x = (double)si;
      }
    }
#undef LB_FRHS
    if (status == 0 && C.trace && n_pushed > C.capacity) status = LB_TRAJ_CAPACITY;
    if constexpr (EV) {
      if (C.out_event_t && tid == 0) C.out_event_t[traj] = ev_t;
    }
    write_final<R>(P, C, S, tid, W, traj, y_r, y_i, status, fail_x, integral, accepted, rhs_calls, n_pushed);
  }
}

// ---------------------------------------------------------------------------
// Paired integrator: two trajectories per wave.
//
// For plans of at most 32 elements each half-wave (32 lanes) integrates its
// own trajectory, so one wave's instruction stream -- the expression machine
// above all, which runs in six lanes per step whatever the plan size --
// serves two trajectories.  The arithmetic is that of the single-wave form:
// the same gather order, stage combinations and controller, and half-wave
// sums built from the same DPP rows as the full-wave sum plus the exact zeros
// the other half contributes there, so results are bitwise those of every
// other variant.  Each half is its own state machine: when its trajectory
// ends it writes the outputs, takes the next one from the queue and
// initialises it while the other half keeps stepping.  Lanes of one wave
// order their LDS traffic with wavefront-scope fences (no workgroup barrier
// inside the divergent halves).
//
// Covered: persistent launches with pushes at every accepted step (final
// outputs and the weighted integral), a given initial step dt, and the
// adjoint's attempt tape; trace, save plans and replays take the other
// kernels.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over this lane's half-wave, rounded as wave_sum rounds a wave whose
// other half holds zeros: the 16-lane rows by the same DPP tree, then
// (row_a + row_b) + (0 + 0).
__device__ __forceinline__ double half_sum(double x, int lane) {
  x += dpp_d<0xB1>(x);
  x += dpp_d<0x4E>(x);
  x += dpp_d<0x141>(x);
  x += dpp_d<0x140>(x);
  const double lo = lane_d(x, 0) + lane_d(x, 16), hi = lane_d(x, 32) + lane_d(x, 48);
  return (lane < 32 ? lo : hi) + (0.0 + 0.0);
}

// A block total of the single-wave form (one block): the stored sum is the
// only non-zero lane of a second wave sum.
__device__ __forceinline__ double half_total(double x, int lane) { return half_sum(x, lane) + 0.0; }

// The single-wave RHS for the element in slot `sl` of this half: inputs from
// the lanes of the same half.  A kernel compiled for one plan (LB_GEN_COEF)
// holds exactly its rows, so its loops have no row checks; the words are
// loaded first, then every permute, then the factors and coefficients, so
// each group of loads issues together.
template <int G>
__device__ __forceinline__ void rhs_pair(const uint32_t* gw, const double2* gf, int rows, int sl, int base,
                                         double2 selfv, bool diag, double xr, double xi, double& kr, double& ki,
                                         const dcpx* coef) {
#ifdef LB_GEN_COEF
  (void)rows;
  constexpr bool exact = true;
#else
  constexpr bool exact = false;
#endif
  double accr = fma(selfv.x, xr, -(selfv.y * xi));
  double acci = fma(selfv.x, xi, selfv.y * xr);
  uint32_t w[G];
  double2 f[G];
  double sr[G], si[G];
#pragma unroll
  for (int j = 0; j < G; ++j)
    if (exact || j < rows) w[j] = gw[j * 64 + sl];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if (exact || j < rows) {
      const int src = (int)(w[j] & 31u) | base;
      sr[j] = __shfl(xr, src, 64);
      si[j] = __shfl(xi, src, 64);
    }
  }
#pragma unroll
  for (int j = 0; j < G; ++j)
    if (exact || j < rows) f[j] = gf[j * 64 + sl];
#pragma unroll
  for (int j = 0; j < G; ++j)
    if (exact || j < rows) gather_apply(w[j], f[j], coef, sr[j], si[j], accr, acci);
  kr = accr;
  ki = diag ? 0.0 : acci;
}

struct LdsPair {
  double2* gf;    // g_rows x 64 gather factors
  dcpx* coef;     // 2 halves x 8 rows x (K + 1): [0] t0, [2..7] stage times
  dcpx* slots;    // 2 halves x 8 x n_slots
  dcpx* vm;       // 2 halves x 8 lanes x vm_depth
  double2* selfc; // 2 halves x 32 folded self coefficients
  double2* park;  // 2 x 64: y and k1 around the expression machine
  double* wts;    // n integral weights
  uint32_t* gw;   // g_rows x 64 gather words
  int* misc;      // [0..1] queue values, [2..3] support flags
};

__device__ __forceinline__ LdsPair carve_pair(const DevPlan& P) {
  extern __shared__ double smem[];
  const int g = cst(P.blk_rows)[0];
  char* q = reinterpret_cast<char*>(smem);
  LdsPair L;
  L.gf = reinterpret_cast<double2*>(q); q += (size_t)g * 64 * 16;
  L.coef = reinterpret_cast<dcpx*>(q); q += (size_t)2 * 8 * (P.K + 1) * 16;
  L.slots = reinterpret_cast<dcpx*>(q); q += (size_t)2 * 8 * P.n_slots * 16;
  L.vm = reinterpret_cast<dcpx*>(q); q += (size_t)2 * 8 * P.vm_depth * 16;
  L.selfc = reinterpret_cast<double2*>(q); q += (size_t)64 * 16;
  L.park = reinterpret_cast<double2*>(q); q += (size_t)128 * 16;
  L.wts = reinterpret_cast<double*>(q); q += (size_t)((P.n + 1) & ~1) * 8;
  L.gw = reinterpret_cast<uint32_t*>(q); q += (size_t)g * 64 * 4;
  L.misc = reinterpret_cast<int*>(q);
  return L;
}

}  // namespace

#ifdef LB_GEN_COEF
// The plan's time-dependent compound and coefficient programs as straight-line
// code: the operations of eval_coefficients(.., only_dynamic = true, ..) in the
// same order on the same operands, so the values are bitwise those of the
// expression machine.  Defined by the source that lb_rtc.cpp generates per
// plan and compiles with this file.
__device__ __forceinline__ void lb_gen_coef(dcpx* sl, dcpx* cf, double t);
#endif

template <int SOLVER, int G, int WAVES>
__global__ __launch_bounds__(64, WAVES) void lb_solve_pair_kernel(const DevPlan P, const DevCall C) {
  using T = Tab<SOLVER>;
  const int lane = threadIdx.x, half = lane >> 5, sl = lane & 31, base = lane & 32;
  const LdsPair L = carve_pair(P);
  const int KK = P.K + 1;
  const int grows = cst(P.blk_rows)[0];
  dcpx* const coef = L.coef + half * 8 * KK;
  dcpx* const slots = L.slots + half * 8 * P.n_slots;
  Slots<1> S;
  load_slots<1>(P, S, sl, 64);
  const bool diag = S.diag(0);
  for (int i = lane; i < grows * 64; i += 64) {
    L.gf[i] = reinterpret_cast<const double2*>(P.g_f)[i];
    L.gw[i] = P.g_word[i];
  }
  for (int i = lane; i < P.n; i += 64) {
    double w = 0.0;
    for (int q = 0; q < C.n_weights; ++q)
      if (C.weight_index[q] == i) w += C.weight_value[q];
    L.wts[i] = w;
  }
  wave_sync();
  const double my_w = (S.valid(0) && diag) ? L.wts[S.state(0)] : 0.0;

  // per-half trajectory state
  bool active = false, done = false;
  int64_t traj = 0;
  double t1 = 0.0, x = 0.0, h = 0.0, hmax = 0.0, fac_old = 1.0e-4;
  double integral = 0.0, last_t = 0.0, last_v = 0.0, fail_x = 0.0;
  int64_t accepted = 0, rejected = 0, rhs_calls = 0, ns = 0, nst = 0, ia = 0;
  int status = 0;
  bool reject = false, last = false;
  double2 selfv = double2{0.0, 0.0};
  double y_r[1] = {0.0}, y_i[1] = {0.0};
  double k_r[8][1], k_i[8][1];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    k_r[j][0] = 0.0;
    k_i[j][0] = 0.0;
  }
  const double beta = 0.04, alpha = 0.2 - beta * 0.75, facc1 = 1.0 / 0.2, facc2 = 1.0 / 10.0;
  const double safety = 0.9, nanv = __builtin_nan("");

  for (;;) {
    // ---- halves without a trajectory take the next one
    const bool need = !active && !done;
    if (need && sl == 0) L.misc[half] = atomicAdd(C.queue, 1);
    wave_sync();
    if (need) {
      const int64_t local = L.misc[half];
      if (local >= C.n_traj) {
        done = true;
      } else {
        traj = C.queue_order ? (int64_t)C.queue_order[local + C.traj_offset] : local + C.traj_offset;
        active = traj >= 0 && traj < C.n_total;  // entries outside the batch are skipped
      }
    }
    if (__all(done)) break;
    // ---- initialise a new trajectory (rust/src/ode/dopri5.rs:160-220)
    const bool starting = need && active;
    if (starting) {
      t1 = C.t1_per_traj ? C.t1_per_traj[traj] : C.t1;
      for (int i = sl; i < P.n_slots; i += 32)
        slots[i] = i < P.n_base ? mk(P.base[2 * i], P.base[2 * i + 1]) : mk(0.0, 0.0);
      wave_sync();
      if (sl == 0) {
        apply_overrides(C, traj, slots);
        eval_coefficients(P, slots, coef, C.t0, false, L.vm + half * 8 * P.vm_depth);
      }
      wave_sync();
      for (int i = sl; i < 7 * P.n_slots; i += 32) slots[P.n_slots + i] = slots[i % P.n_slots];
      for (int i = sl; i < 7 * KK; i += 32) coef[KK + i] = coef[i % KK];
      wave_sync();  // the stage rows are read by the expression-machine lanes
      {  // folded self coefficient of this element (fold_self arithmetic)
        dcpx acc = mk(0.0, 0.0);
        for (int j = 0; j < P.max_self; ++j) {
          const size_t e = (size_t)sl * P.max_self + j;
          const int id = P.self_coef[e];
          if (id == -2) break;
          const dcpx f = mk(P.self_f[2 * e], P.self_f[2 * e + 1]);
          dcpx c = id < 0 ? mk(1.0, 0.0) : coef[id];
          if (P.self_conj[e]) c.im = -c.im;
          acc = cadd(acc, cmul(f, c));
        }
        selfv = double2{acc.re, acc.im};
      }
      const double* y0 = C.y0 + (C.y0_shared ? 0 : traj * (int64_t)P.dim);
      y_r[0] = S.valid(0) ? y0[P.slot_re[sl]] : 0.0;
      y_i[0] = (S.valid(0) && !diag) ? y0[P.slot_im[sl]] : 0.0;
      bool bad_support = false;
      if (P.restricted) {
        if (sl == 0) L.misc[2 + half] = 0;
        wave_sync();
        for (int k = sl; k < P.dim; k += 32)
          if (!P.upper_active[P.packed_upper[k]] && y0[k] != 0.0) L.misc[2 + half] = 1;
        wave_sync();
        bad_support = L.misc[2 + half] != 0;
      }
      accepted = rejected = rhs_calls = 0;
      status = 0;
      fail_x = 0.0;
      integral = 0.0;
      last_t = __builtin_nan("");
      last_v = 0.0;
      x = C.t0;
      reject = false;
      last = false;
      fac_old = 1.0e-4;
      ns = nst = ia = 0;
      hmax = fabs(t1 - C.t0);
      if (bad_support) {
        status = LB_TRAJ_BAD_SUPPORT;
        fail_x = C.t0;
        last = true;
      } else if (t1 < C.t0) {
        status = LB_TRAJ_BAD_SPAN;
        fail_x = C.t0;
        last = true;
      } else {
        // pushed point t0 (save_start): the trapezoid's first node
        if (C.save_start && C.want_integral) {
          last_v = half_total(S.valid(0) && diag ? my_w * y_r[0] : 0.0, lane);
          last_t = C.t0;
        }
        if (t1 == C.t0) {
          last = true;
        } else {
          rhs_pair<G>(L.gw, L.gf, grows, sl, base, selfv, diag, y_r[0], y_i[0], k_r[1][0], k_i[1][0], coef);
          h = fmin(C.dt, t1 - C.t0);
          rhs_calls = 1;
        }
      }
    }
    // ---- one step attempt of every half that is integrating
    const bool stepping = active && !last;
    if (stepping) {
      if (ns > C.maxiters) {
        status = LB_TRAJ_MAXITERS;
        fail_x = x;
        last = true;
      } else if (0.1 * fabs(h) <= DBL_EPSILON * fabs(x)) {
        status = LB_TRAJ_UNDERFLOW;
        fail_x = x;
        last = true;
      }
    }
    const bool attempt = stepping && !last;
    bool finishing = false;
    if (attempt) {
      if (x + 1.01 * h - t1 > 0.0) {
        h = t1 - x;
        finishing = true;
      }
      ++ns;
      L.park[lane] = double2{y_r[0], y_i[0]};
      L.park[64 + lane] = double2{k_r[1][0], k_i[1][0]};
      if (P.time_dependent && sl < 6) {
        const int st = sl + 2;
#ifdef LB_GEN_COEF
        lb_gen_coef(slots + st * P.n_slots, coef + st * KK, x + h * T::c(st));
#else
        eval_coefficients(P, slots + st * P.n_slots, coef + st * KK, x + h * T::c(st), true,
                          L.vm + (half * 8 + sl) * P.vm_depth);
#endif
      }
      wave_sync();  // stage coefficients are read by every lane of the half
      {
        const double2 py = L.park[lane], pk = L.park[64 + lane];
        y_r[0] = py.x;
        y_i[0] = py.y;
        k_r[1][0] = pk.x;
        k_i[1][0] = pk.y;
      }
      const bool need_stiff = ((accepted + 1) % 1000 == 0) || ia > 0;
      double yt_r[1], yt_i[1], yn_r[1], yn_i[1];
#define LB_PAIR_STAGE(st)                                                                            \
  stage_combo<1, SOLVER, st>(h, y_r, y_i, k_r, k_i, yt_r, yt_i);                                    \
  rhs_pair<G>(L.gw, L.gf, grows, sl, base, selfv, diag, yt_r[0], yt_i[0], k_r[st][0], k_i[st][0],   \
              coef + (st) * KK);
      LB_PAIR_STAGE(2)
      LB_PAIR_STAGE(3)
      LB_PAIR_STAGE(4)
      LB_PAIR_STAGE(5)
      LB_PAIR_STAGE(6)
#undef LB_PAIR_STAGE
      stage_combo<1, SOLVER, 7>(h, y_r, y_i, k_r, k_i, yn_r, yn_i);
      double stiff_d = 0.0;
      if (need_stiff) {
        const double a0 = yn_r[0] - yt_r[0], a1 = yn_i[0] - yt_i[0];
        stiff_d = a0 * a0 + a1 * a1;
      }
      rhs_pair<G>(L.gw, L.gf, grows, sl, base, selfv, diag, yn_r[0], yn_i[0], k_r[7][0], k_i[7][0], coef + 7 * KK);
      rhs_calls += 6;
      // ---- error norm, integrand, stiffness (as the single-wave form)
      double er = T::e(1) * k_r[1][0], ei = T::e(1) * k_i[1][0];
#pragma unroll
      for (int j = 2; j <= 7; ++j) {
        if (T::e(j) != 0.0) {
          er = fma(T::e(j), k_r[j][0], er);
          ei = fma(T::e(j), k_i[j][0], ei);
        }
      }
      er *= h;
      ei *= h;
      const double sr = C.abstol + fmax(fabs(y_r[0]), fabs(yn_r[0])) * C.reltol;
      const double si = C.abstol + fmax(fabs(y_i[0]), fabs(yn_i[0])) * C.reltol;
      const double qr = er / sr, qi = ei / si;
      const double e2 = half_total(S.valid(0) ? qr * qr + (diag ? 0.0 : qi * qi) : 0.0, lane);
      const double err = sqrt(e2 / (double)P.dim);
      // the controller's two powers at once: lanes 0-15 of the half take
      // err^alpha, lanes 16-31 fac_old^-beta (the same pow on the same inputs)
      const double pw = pow(sl < 16 ? err : fac_old, sl < 16 ? alpha : -beta);
      const double fac11 = __shfl(pw, base, 64), fo = __shfl(pw, base + 16, 64);
      double fac = fac11 * fo;
      const double fac_s = fac / safety;
      fac = fmax(facc2, fmin(facc1, fac_s));
      double hn = h / fac;
      if (C.tape_h) {
        // attempt tape, as the single-wave form records it: one record per
        // attempt by lane 0 of the half, the state every ckpt_every-th attempt
        const int64_t a = ns - 1;
        if (a < C.tape_cap) {
          if (sl == 0) {
            int fl = finishing ? LB_TAPE_CLIPPED : 0;
            double h_next;
            if (err <= 1.0) {
              h_next = hn;
              fl |= LB_TAPE_ACCEPTED | (accepted == 0 ? LB_TAPE_FIRST : 0) | (err > 1.0e-4 ? LB_TAPE_ERRQ : 0);
              if (!(fac_s > facc2 && fac_s < facc1)) fl |= LB_TAPE_CLAMPED;
              double hc = hn;
              if (fabs(hc) > hmax) {
                hc = hmax;
                fl |= LB_TAPE_HMAX;
              }
              if (reject && fabs(h) < fabs(hc)) fl |= LB_TAPE_MINRULE;
            } else {
              const double fr = fac11 / safety;
              h_next = h / fmin(facc1, fr);
              if (!(fr < facc1)) fl |= LB_TAPE_CLAMPED;
            }
            const int64_t o = traj * C.tape_cap + a;
            C.tape_t[o] = x;
            C.tape_h[o] = h;
            C.tape_c[3 * o] = err;
            C.tape_c[3 * o + 1] = h_next;
            C.tape_c[3 * o + 2] = fac_old;
            C.tape_f[o] = fl;
          }
          if (a % C.ckpt_every == 0 && sl < P.S) {
            const int64_t ncap = (C.tape_cap + C.ckpt_every - 1) / C.ckpt_every;
            double* ck = C.tape_y + ((traj * ncap + a / C.ckpt_every) * P.S) * 2;
            ck[2 * sl] = y_r[0];
            ck[2 * sl + 1] = y_i[0];
          }
        }
      }
      if (err <= 1.0) {
        fac_old = fmax(err, 1.0e-4);
        if (fabs(hn) > hmax) hn = hmax;
        if (reject) hn = fmin(fabs(hn), fabs(h));
        reject = false;
        ++accepted;
        bool stiff = false;
        if (accepted % 1000 == 0 || ia > 0) {
          double a0 = k_r[7][0] - k_r[6][0], a1 = k_i[7][0] - k_i[6][0];
          const double n2 = half_total(a0 * a0 + a1 * a1, lane), dd2 = half_total(stiff_d, lane);
          const double hl = dd2 > 0.0 ? h * sqrt(n2 / dd2) : 0.0;
          if (hl > 3.25) {
            ++ia;
            nst = 0;
            if (ia == 15) {
              status = LB_TRAJ_STIFF;
              fail_x = x;
              stiff = true;
            }
          } else {
            ++nst;
            if (nst == 6) ia = 0;
          }
        }
        if (stiff) {
          last = true;
        } else {
          x += h;
          if (C.want_integral) {
            const double v = half_total(S.valid(0) && diag ? my_w * yn_r[0] : 0.0, lane);
            if (isfinite(last_t)) integral += 0.5 * (last_v + v) * (x - last_t);
            last_t = x;
            last_v = v;
          }
          y_r[0] = yn_r[0];
          y_i[0] = yn_i[0];
          k_r[1][0] = k_r[7][0];
          k_i[1][0] = k_i[7][0];
          if (finishing) last = true;
          h = hn;
        }
      } else {
        hn = h / fmin(facc1, fac11 / safety);
        reject = true;
        if (accepted >= 1) ++rejected;
        h = hn;
      }
    }
    // ---- a finished trajectory writes its outputs (the single-wave form's)
    if (active && last) {
      if (status == 0 && C.tape_h && ns > C.tape_cap) status = LB_TRAJ_CAPACITY;
      const bool ok = status == 0;
      if (P.restricted) {
        if (C.output_kind == LB_OUT_POPULATIONS) {
          for (int i = sl; i < P.n; i += 32)
            if (!P.upper_active[P.packed_upper[i]]) C.out_values[traj * P.n + i] = ok ? 0.0 : nanv;
        } else if (C.output_kind == LB_OUT_FULL) {
          for (int k = sl; k < P.dim; k += 32)
            if (!P.upper_active[P.packed_upper[k]]) C.out_values[traj * P.dim + k] = ok ? 0.0 : nanv;
        } else if (C.output_kind == LB_OUT_SELECTED) {
          for (int q = sl; q < C.n_selected; q += 32)
            if (!P.upper_active[C.sel_upper[q]]) {
              C.out_values[2 * (traj * C.n_selected + q)] = ok ? 0.0 : nanv;
              C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? 0.0 : nanv;
            }
        }
      }
      if (S.valid(0)) {
        if (C.output_kind == LB_OUT_POPULATIONS) {
          if (diag) C.out_values[traj * P.n + S.state(0)] = ok ? y_r[0] : nanv;
        } else if (C.output_kind == LB_OUT_FULL) {
          C.out_values[traj * P.dim + P.slot_re[sl]] = ok ? y_r[0] : nanv;
          if (!diag) C.out_values[traj * P.dim + P.slot_im[sl]] = ok ? y_i[0] : nanv;
        } else if (C.output_kind == LB_OUT_SELECTED) {
          const int up = P.slot_up[sl];
          for (int q = 0; q < C.n_selected; ++q) {
            if (C.sel_upper[q] != up) continue;
            const double im = (C.sel_flags[q] & 1) ? 0.0 : ((C.sel_flags[q] & 2) ? -y_i[0] : y_i[0]);
            C.out_values[2 * (traj * C.n_selected + q)] = ok ? y_r[0] : nanv;
            C.out_values[2 * (traj * C.n_selected + q) + 1] = ok ? im : nanv;
          }
        }
      }
      if (sl == 0) {
        if (C.tape_n) C.tape_n[traj] = (int32_t)ns;
        if (C.output_kind == LB_OUT_INTEGRAL) C.out_values[traj] = ok ? integral : nanv;
        if (C.out_integral) C.out_integral[traj] = ok ? integral : nanv;
        C.out_stats[4 * traj + 0] = accepted;
        C.out_stats[4 * traj + 1] = rejected;
        C.out_stats[4 * traj + 2] = rhs_calls;
        C.out_stats[4 * traj + 3] = status;
        C.out_fail_t[traj] = fail_x;
      }
      active = false;
    }
  }
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the adaptive Dopri5/Tsit5 solve.
//
// The objective is the weighted trapezoid I over the pushed points (t0 when
// saving the start, then every accepted step; rust/src/ode/output.rs:315-330).
// The forward solve recorded an attempt tape: for every step attempt its
// start x, size h, error norm, the controller's proposed next step and
// fac_old, and its flags (accepted, clipped to t1, which controller branch),
// plus the state at the start of every ckpt_every-th attempt.  Segments of
// ckpt_every attempts are processed last to first: the segment is recomputed
// from its checkpoint, keeping each attempt's start state y and its stage
// derivatives k1..k7 in the workgroup's workspace, then its attempts are
// swept backwards.  One attempt maps (y, x, h, fac_old) to the candidate
// yn = y + h sum_j b_j k_j, the error estimate e = h sum_j e_j k_j and, through
// err = ||e / sc||_rms with sc = atol + rtol max(|y|, |yn|) (dopri5.rs:88-102,
// tsit5.rs:104-119), to the next proposed step (common.rs:48-69).  With
// lambda the cotangent of yn and errb that of err, nu = errb de/derr,
//   kappa_7 = h e_7 nu,  mu_7 = A(x + h)^T kappa_7,  lambda' = lambda + mu_7,
//   kappa_i = h (b_i lambda' + e_i nu + sum_{j>i} a_ji mu_j),  mu_i = A(t_i)^T kappa_i,
//   ybar = lambda' + sum_i mu_i (+ the sc terms), dI/dp += sum_i kappa_i^T dA(t_i)/dp Y_i,
// and the scalar cotangents of x and h collect kappa_i^T dA(t_i)/dt Y_i (time
// derivatives of the coefficient programs) and the explicit h terms.  The
// scalar chain -- the controller's partial derivatives, the trapezoid's
// dependence on h, x_{n+1} = x_n + h, the clip h = t1 - x -- links the
// attempts.  A^T comes from the plan's transposed term tables and dA/dp,
// dA/dt from the derivative coefficient programs (the RHS is linear in rho).
// LB_ADJOINT_FROZEN keeps only the accepted attempts with their step sizes
// fixed (the derivative of a replay).
// ---------------------------------------------------------------------------
namespace {

// d h_next / d (h, err, fac_old) of one attempt (common.rs:48-69): interior
// branches h_next = h safety fac_old^beta / err^alpha (accepted) or
// h safety / err^alpha (rejected); a clamped factor leaves h / facc; the h_max
// cut makes it constant and the after-reject rule makes it h.
__device__ __forceinline__ void ctrl_partials(int fl, double h, double err, double h_next, double q, double& dh,
                                              double& de, double& dq) {
  const double beta = 0.04, alpha = 0.2 - beta * 0.75;
  const bool acc = (fl & LB_TAPE_ACCEPTED) != 0;
  dh = h_next / h;
  de = 0.0;
  dq = 0.0;
  if (!(fl & LB_TAPE_CLAMPED)) {
    de = -alpha * h_next / err;
    if (acc) dq = beta * h_next / q;
  }
  if (acc && (fl & LB_TAPE_MINRULE)) {
    dh = 1.0;
    de = 0.0;
    dq = 0.0;
  } else if (acc && (fl & LB_TAPE_HMAX)) {
    dh = 0.0;
    de = 0.0;
    dq = 0.0;
  }
}

// The error-norm cotangents of one packed component: nu = errb d err / d e,
// and the scale's term, which goes to the start state or the candidate
// (whichever max(|y|, |yn|) took).
__device__ __forceinline__ void err_component(double errb, double err, double dimf, double atol, double rtol,
                                              double e, double y, double yn, double& nu, double& yb, double& lb) {
  const double sc = atol + fmax(fabs(y), fabs(yn)) * rtol;
  const double q = e / sc, den = dimf * err * sc;
  nu = errb * q / den;
  const double scb = -errb * q * q / den * rtol;
  if (fabs(y) >= fabs(yn)) yb += y > 0.0 ? scb : (y < 0.0 ? -scb : 0.0);
  else lb += yn > 0.0 ? scb : (yn < 0.0 ? -scb : 0.0);
}

// Per-lane pieces of dI/dh that are inner products with the stage
// derivatives: lambda'.sum b_j k_j + nu.sum e_j k_j + sum_i mu_i.sum_{j<i} a_ij k_j.
template <int SOLVER>
__device__ __forceinline__ double h_products(const double (&k_r)[8][1], const double (&k_i)[8][1], double lpr,
                                             double lpi, double nur, double nui, const double (&mu_r)[8],
                                             const double (&mu_i)[8]) {
  using T = Tab<SOLVER>;
  double sbr = 0.0, sbi = 0.0, ser = 0.0, sei = 0.0;
#pragma unroll
  for (int j = 1; j <= 7; ++j) {
    if (j <= 6) {
      sbr += T::a(7, j) * k_r[j][0];
      sbi += T::a(7, j) * k_i[j][0];
    }
    ser += T::e(j) * k_r[j][0];
    sei += T::e(j) * k_i[j][0];
  }
  double v = lpr * sbr + lpi * sbi + nur * ser + nui * sei;
#pragma unroll
  for (int i = 2; i <= 6; ++i) {
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int j = 1; j < i; ++j) {
      sr += T::a(i, j) * k_r[j][0];
      si += T::a(i, j) * k_i[j][0];
    }
    v += mu_r[i] * sr + mu_i[i] * si;
  }
  return v;
}

struct LdsAdj {
  double* pub;    // 2*n_pub: published stage input Y_i (forward publication indices)
  double* kap;    // 2*S: published kappa_i by slot
  double* selfc;  // 2*S folded self coefficients
  double* wts;    // n integral weights per state
  dcpx* coef;     // 1 + 7 E rows x (KD + 1): [0] t0, then the stage times of E attempts
  dcpx* slots;    // 1 + 7 E slot copies (n_slots each)
  double* red;    // LB_MAX_DIFF x 16 per-wave sums
  dcpx* vm;       // 8 lanes x vm_depth
  int* misc;
};

__device__ __forceinline__ LdsAdj carve_adj(const DevPlan& P, int E) {
  extern __shared__ double smem[];
  const size_t rows = 1 + 7 * (size_t)E;
  LdsAdj L;
  L.pub = smem;
  L.kap = L.pub + 2 * (size_t)P.n_pub;
  L.selfc = L.kap + 2 * (size_t)P.S;
  L.wts = L.selfc + 2 * (size_t)P.S;
  L.coef = reinterpret_cast<dcpx*>(L.wts + ((P.n + 1) & ~1));
  L.slots = L.coef + rows * (P.KD + 1);
  L.red = reinterpret_cast<double*>(L.slots + rows * P.n_slots);
  L.vm = reinterpret_cast<dcpx*>(L.red + LB_MAX_DIFF * 16);
  L.misc = reinterpret_cast<int*>(L.vm + max(8, 7 * E) * (size_t)P.vm_depth);
  return L;
}

// Forward RHS for the recomputation: publish, barrier, self term + gathers
// (tables read from global memory), barrier (the buffer is reused).
template <int R>
__device__ __forceinline__ void rhs_adj_fwd(const DevPlan& P, const LdsAdj& L, const Slots<R>& S,
                                            const double (&xr)[R], const double (&xi)[R], double (&kr)[R],
                                            double (&ki)[R], double* pubbuf, const dcpx* coef, int tid, int W) {
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (S.pub[r] >= 0) {
      pubbuf[2 * S.pub[r]] = xr[r];
      pubbuf[2 * S.pub[r] + 1] = xi[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int blk = wave + (W >> 6) * r;
    if (blk >= P.nblk) { kr[r] = 0.0; ki[r] = 0.0; continue; }
    const int s = tid + W * r;
    const double2 a = reinterpret_cast<const double2*>(L.selfc)[s];
    double accr = fma(a.x, xr[r], -(a.y * xi[r]));
    double acci = fma(a.x, xi[r], a.y * xr[r]);
    gather_block(P.g_word, reinterpret_cast<const double2*>(P.g_f), cst(P.blk_row0)[blk], cst(P.blk_rows)[blk],
                 lane, coef, pubbuf, accr, acci);
    kr[r] = accr;
    ki[r] = S.diag(r) ? 0.0 : acci;
  }
  __syncthreads();
}

// mu = A(t)^T kappa for the slots of this thread, the parameter contraction
// g_m += kappa^T (dA/dp_m) Y and (plans with a time column) the time
// contraction tc += kappa^T (dA/dt) Y over the terms of its outputs.  kappa
// and Y are published before the call (kap by slot, pub by publication index)
// and a barrier separates the publication from this call.
template <int R>
__device__ __forceinline__ void adjoint_stage(const DevPlan& P, const LdsAdj& L, const Slots<R>& S,
                                              const double (&kr)[R], const double (&ki)[R],
                                              const double (&yr)[R], const double (&yi)[R],
                                              const dcpx* cf, double (&mr)[R], double (&mi)[R],
                                              double (&g)[LB_MAX_DIFF], double& tc, int tid, int W) {
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nd = P.n_diff, np = P.n_diff - P.diff_time;
  const dcpx* dc = cf + P.K + 1;  // dc[k * nd + m] = d c_k / d p_m (m = np: d c_k / d t)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    mr[r] = 0.0;
    mi[r] = 0.0;
    const int blk = wave + (W >> 6) * r;
    if (blk >= P.nblk) continue;
    const int s = tid + W * r;
    const bool diag = S.diag(r);
    // conj(kappa_own)
    const double ckr = kr[r], cki = -ki[r];
    // self terms: transposed (conj(kappa) selfc) and their parameter derivatives
    // (their coefficients do not depend on time)
    const double2 a = reinterpret_cast<const double2*>(L.selfc)[s];
    {
      const double hr = ckr * a.x - cki * a.y, hi = ckr * a.y + cki * a.x;
      mr[r] += hr;
      if (!diag) mi[r] -= hi;
    }
    if (np > 0) {
      for (int j = 0; j < P.max_self; ++j) {
        const size_t e = (size_t)s * P.max_self + j;
        const int id = P.self_coef[e];
        if (id == -2) break;
        if (id < 0) continue;
        const dcpx f = mk(P.self_f[2 * e], P.self_f[2 * e + 1]);
        // Q = conj(kappa) f x_own
        const dcpx x = mk(yr[r], diag ? 0.0 : yi[r]);
        const dcpx q = cmul(mk(ckr, cki), cmul(f, x));
        const bool cj = P.self_conj[e] != 0;
        for (int m = 0; m < LB_MAX_DIFF; ++m) {
          if (m >= np) break;
          const dcpx d = dc[id * nd + m];
          g[m] += q.re * d.re - q.im * (cj ? -d.im : d.im);
        }
      }
    }
    // transposed gathers: outputs o reading this slot
    const int trows = cst(P.t_blk_rows)[blk], trow0 = cst(P.t_blk_row0)[blk];
    for (int j = 0; j < trows; ++j) {
      const size_t idx = (size_t)(trow0 + j) * 64 + lane;
      const uint32_t w = P.t_word[idx];
      const double2 f = reinterpret_cast<const double2*>(P.t_f)[idx];
      const uint32_t o = w & 0xFFFFFu;
      const uint32_t cid = w >> 24;
      const dcpx cl = cf[cid == LB_STATIC_COEF ? 0u : cid];
      dcpx c = cid == LB_STATIC_COEF ? mk(1.0, 0.0) : cl;
      if (w & (1u << 22)) c.im = -c.im;
      // H = conj(kappa_o) f C
      const dcpx ko = mk(L.kap[2 * o], -L.kap[2 * o + 1]);
      const dcpx hh = cmul(ko, cmul(mk(f.x, f.y), c));
      mr[r] += hh.re;
      if (!(w & (1u << 20))) mi[r] += (w & (1u << 21)) ? hh.im : -hh.im;
    }
    // coefficient derivatives of the gather terms of this output
    if (nd > 0) {
      const int rows = cst(P.blk_rows)[blk], row0 = cst(P.blk_row0)[blk];
      for (int j = 0; j < rows; ++j) {
        const size_t idx = (size_t)(row0 + j) * 64 + lane;
        const uint32_t w = P.g_word[idx];
        const uint32_t cid = w >> 24;
        if (cid == LB_STATIC_COEF) continue;
        const double2 f = reinterpret_cast<const double2*>(P.g_f)[idx];
        const uint32_t p = w & 0xFFFFFu;
        double xr = L.pub[2 * p];
        double xi = (w & (1u << 20)) ? 0.0 : L.pub[2 * p + 1];
        if (w & (1u << 21)) xi = -xi;
        const dcpx q = cmul(mk(ckr, cki), cmul(mk(f.x, f.y), mk(xr, xi)));
        const bool cj = (w & (1u << 22)) != 0;
        for (int m = 0; m < LB_MAX_DIFF; ++m) {
          if (m >= np) break;
          const dcpx d = dc[cid * nd + m];
          g[m] += q.re * d.re - q.im * (cj ? -d.im : d.im);
        }
        if (P.diff_time) {
          const dcpx d = dc[cid * nd + np];
          tc += q.re * d.re - q.im * (cj ? -d.im : d.im);
        }
      }
    }
  }
}

// Sum of one value over the workgroup (every thread gets the same value): a
// wave sum, then the waves in index order.  Synchronises.
__device__ __forceinline__ double group_sum(double v, double* red, int tid, int W) {
  const double s = wave_sum(v);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int b = 0; b < (W >> 6); ++b) t += red[b];
  __syncthreads();
  return t;
}

}  // namespace

template <int R, int SOLVER, int MAXW, int WAVES>
__global__ __launch_bounds__(MAXW, WAVES) void lb_adjoint_kernel(const DevPlan P, const DevCall C) {
  using T = Tab<SOLVER>;
  const int tid = threadIdx.x, W = blockDim.x;
  // the coefficient rows of E attempts are evaluated in one pass (7 E lanes)
  const int E = max(1, C.adj_row_batch);
  const LdsAdj L = carve_adj(P, E);
  const int KC = P.KD + 1, np = P.n_diff - P.diff_time;
  const bool adaptive = C.adj_mode == LB_ADJOINT_ADAPTIVE;
  const double dimf = (double)P.dim;
  Slots<R> S;
  load_slots<R>(P, S, tid, W);
  const int every = C.ckpt_every;
  // workspace per attempt: y and k1..k7 (8 x S complex), 7 coefficient rows
  double* segY = C.seg_y + (size_t)blockIdx.x * every * 8 * P.S * 2;
  dcpx* segC = reinterpret_cast<dcpx*>(C.seg_c) + (size_t)blockIdx.x * every * 7 * KC;
  const int64_t ncap = (C.tape_cap + every - 1) / every;
  auto ws = [&](int a, int j, int s) -> double* { return segY + ((size_t)(a * 8 + j) * P.S + s) * 2; };

  for (;;) {
    __syncthreads();
    if (tid == 0) L.misc[0] = atomicAdd(C.queue, 1);
    __syncthreads();
    const int64_t local = L.misc[0];
    if (local >= C.n_traj) break;
    // queue position -> trajectory (longest-first orders); entries outside
    // the batch are skipped, so a bad order cannot address past the arrays
    const int64_t traj = C.queue_order ? (int64_t)C.queue_order[local + C.traj_offset] : local + C.traj_offset;
    if (traj < 0 || traj >= C.n_total) continue;

    // ---- parameters, coefficients at t0 (self folds), weights
    for (int i = tid; i < P.n_slots; i += W)
      L.slots[i] = i < P.n_base ? mk(P.base[2 * i], P.base[2 * i + 1]) : mk(0.0, 0.0);
    __syncthreads();
    if (tid == 0) {
      apply_overrides(C, traj, L.slots);
      eval_coefficients(P, L.slots, L.coef, C.t0, false, L.vm, P.KD);
    }
    for (int i = tid; i < P.n; i += W) {
      double w = 0.0;
      for (int q = 0; q < C.n_weights; ++q)
        if (C.weight_index[q] == i) w += C.weight_value[q];
      L.wts[i] = w;
    }
    __syncthreads();
    // stage rows start as the t0 row: the expression machine rewrites only the
    // time-dependent entries at each stage time
    for (int i = tid; i < 7 * E * P.n_slots; i += W) L.slots[P.n_slots + i] = L.slots[i % P.n_slots];
    for (int i = tid; i < 7 * E * KC; i += W) L.coef[KC + i] = L.coef[i % KC];
    fold_self(P, tid, W, L.coef, L.selfc);
    __syncthreads();

    double g[LB_MAX_DIFF];
#pragma unroll
    for (int m = 0; m < LB_MAX_DIFF; ++m) g[m] = 0.0;
    const int N = C.tape_n[traj];
    const bool ok = N > 0 && N <= C.tape_cap;
    if (ok) {
      const double* tt = C.tape_t + traj * C.tape_cap;
      const double* th = C.tape_h + traj * C.tape_cap;
      const double* tcr = C.tape_c + traj * C.tape_cap * 3;
      const int32_t* tfl = C.tape_f + traj * C.tape_cap;
      double wme[R];
#pragma unroll
      for (int r = 0; r < R; ++r) wme[r] = (S.valid(r) && S.diag(r)) ? L.wts[S.state(r)] : 0.0;
      // cotangents: lam of the state at the start of the current step; Tb of
      // its start time; Hb of the step size proposed into the attempt after
      // the current one (before its clip); Qb of fac_old after the current
      // step
      double lam_r[R], lam_i[R];
#pragma unroll
      for (int r = 0; r < R; ++r) { lam_r[r] = 0.0; lam_i[r] = 0.0; }
      double Tb = 0.0, Hb = 0.0, Qb = 0.0;
      const int nseg = (N + every - 1) / every;
      for (int sg = nseg - 1; sg >= 0; --sg) {
        const int a0 = sg * every, a1 = min(a0 + every, N);
        // ---- recompute the segment from its checkpoint
        double y_r[R], y_i[R];
        {
          const double* ck = C.tape_y + ((traj * ncap + sg) * P.S) * 2;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int s = tid + W * r;
            y_r[r] = s < P.S ? ck[2 * s] : 0.0;
            y_i[r] = s < P.S ? ck[2 * s + 1] : 0.0;
          }
        }
        for (int a = a0; a < a1; ++a) {
          const double h = th[a];
          const int jb = (a - a0) % E;  // this attempt's place in its batch of E
          if (jb == 0) {
            // the coefficient rows of the next E attempts, lane 7 j + s - 1
            // for attempt a + j, stage s; kept for the sweep in segC
            __syncthreads();  // previous users of the coefficient rows are done
            const int nb = min(E, a1 - a);
            if (tid < 7 * nb) {
              const int aa = a + tid / 7, st = tid % 7 + 1;
              const double ts = st == 1 ? tt[aa] : tt[aa] + th[aa] * T::c(st);
              dcpx* const row = L.coef + (1 + tid) * KC;
              eval_coefficients(P, L.slots + (1 + tid) * P.n_slots, row, ts, true, L.vm + tid * P.vm_depth, P.KD);
              for (int q = 0; q < KC; ++q) segC[((aa - a0) * 7 + st - 1) * KC + q] = row[q];
            }
            __syncthreads();
          }
          // stage s of this attempt at rows[s] (row 0 of the block is the t0 row)
          const dcpx* const rows = L.coef + 7 * jb * KC;
          double k_r[8][R], k_i[8][R], yt_r[R], yt_i[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int s = tid + W * r;
            if (s < P.S) {
              double* w0 = ws(a - a0, 0, s);
              w0[0] = y_r[r];
              w0[1] = y_i[r];
            }
          }
          rhs_adj_fwd<R>(P, L, S, y_r, y_i, k_r[1], k_i[1], L.pub, rows + KC, tid, W);
#define LB_ADJ_KEEP(st)                                             \
  _Pragma("unroll") for (int r = 0; r < R; ++r) {                   \
    const int s = tid + W * r;                                      \
    if (s < P.S) {                                                  \
      double* wk = ws(a - a0, st, s);                               \
      wk[0] = k_r[st][r];                                           \
      wk[1] = k_i[st][r];                                           \
    }                                                               \
  }
          LB_ADJ_KEEP(1)
#define LB_ADJ_FWD(st)                                                                              \
  stage_combo<R, SOLVER, st>(h, y_r, y_i, k_r, k_i, yt_r, yt_i);                                   \
  rhs_adj_fwd<R>(P, L, S, yt_r, yt_i, k_r[st], k_i[st], L.pub, rows + (st) * KC, tid, W);          \
  LB_ADJ_KEEP(st)
          LB_ADJ_FWD(2)
          LB_ADJ_FWD(3)
          LB_ADJ_FWD(4)
          LB_ADJ_FWD(5)
          LB_ADJ_FWD(6)
          LB_ADJ_FWD(7)
#undef LB_ADJ_FWD
#undef LB_ADJ_KEEP
          if (tfl[a] & LB_TAPE_ACCEPTED) {
            // stage_combo<7> left the candidate solution in yt
#pragma unroll
            for (int r = 0; r < R; ++r) {
              y_r[r] = yt_r[r];
              y_i[r] = yt_i[r];
            }
          }
        }
        // ---- reverse sweep over the segment's attempts, last first
        for (int a = a1 - 1; a >= a0; --a) {
          const int fl = tfl[a];
          const bool acc = (fl & LB_TAPE_ACCEPTED) != 0;
          if (!adaptive && !acc) continue;
          const double h = th[a];
          double hb = 0.0, errb = 0.0, qnew = Qb;
          if (adaptive) {
            double dh, de, dq;
            ctrl_partials(fl, h, tcr[3 * a], tcr[3 * a + 1], tcr[3 * a + 2], dh, de, dq);
            if (acc) {
              errb = Hb * de + ((fl & LB_TAPE_ERRQ) ? Qb : 0.0);
              hb = Hb * dh + Tb;
              qnew = Hb * dq;
            } else {
              errb = Hb * de;
              hb = Hb * dh;
            }
            // a rejected attempt whose proposal nothing downstream depends on
            if (!acc && errb == 0.0 && hb == 0.0) {
              if (fl & LB_TAPE_CLIPPED) Hb = 0.0;
              continue;
            }
          }
          __syncthreads();
          for (int q = tid; q < 7 * KC; q += W) L.coef[KC + q] = segC[(a - a0) * 7 * KC + q];
          double y_r[R], y_i[R], k_r[8][R], k_i[8][R], yn_r[R], yn_i[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int s = tid + W * r;
            const bool in = s < P.S;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const double* wk = in ? ws(a - a0, j, s) : nullptr;
              const double vr = in ? wk[0] : 0.0, vi = in ? wk[1] : 0.0;
              if (j == 0) { y_r[r] = vr; y_i[r] = vi; } else { k_r[j][r] = vr; k_i[j][r] = vi; }
            }
            k_r[0][r] = 0.0;
            k_i[0][r] = 0.0;
          }
          stage_combo<R, SOLVER, 7>(h, y_r, y_i, k_r, k_i, yn_r, yn_i);
          // lambda (cotangent of the candidate), ybar (of the start state),
          // nu (of the error estimate), per-lane h and t cotangent pieces
          double lr[R], li[R], ybr[R], ybi[R], nur[R], nui[R];
          double hbl = 0.0, tbl = 0.0;
          const bool trap = acc && (!(fl & LB_TAPE_FIRST) || C.save_start);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            lr[r] = acc ? lam_r[r] : 0.0;
            li[r] = acc ? lam_i[r] : 0.0;
            ybr[r] = 0.0;
            ybi[r] = 0.0;
            nur[r] = 0.0;
            nui[r] = 0.0;
            if (trap) {
              // 0.5 (v_n + v_{n+1}) h with v = w . y
              lr[r] += 0.5 * h * wme[r];
              ybr[r] += 0.5 * h * wme[r];
              hbl += 0.5 * wme[r] * (y_r[r] + yn_r[r]);
            }
          }
          if (errb != 0.0) {
            const double err = tcr[3 * a];
#pragma unroll
            for (int r = 0; r < R; ++r) {
              if (!S.valid(r)) continue;
              double er = T::e(1) * k_r[1][r], ei = T::e(1) * k_i[1][r];
#pragma unroll
              for (int j = 2; j <= 7; ++j) {
                if (T::e(j) != 0.0) {
                  er = fma(T::e(j), k_r[j][r], er);
                  ei = fma(T::e(j), k_i[j][r], ei);
                }
              }
              err_component(errb, err, dimf, C.abstol, C.reltol, er * h, y_r[r], yn_r[r], nur[r], ybr[r], lr[r]);
              if (!S.diag(r))
                err_component(errb, err, dimf, C.abstol, C.reltol, ei * h, y_i[r], yn_i[r], nui[r], ybi[r], li[r]);
            }
          }
          double mu_r[8][R], mu_i[8][R];
#pragma unroll
          for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int r = 0; r < R; ++r) { mu_r[i][r] = 0.0; mu_i[i][r] = 0.0; }
          // one reverse stage: publish kappa and the stage input, A^T kappa
#define LB_ADJ_REV(i, KR, KI, YR, YI)                                                              \
  {                                                                                                \
    _Pragma("unroll") for (int r = 0; r < R; ++r) {                                                \
      const int s = tid + W * r;                                                                   \
      if (S.pub[r] >= 0) {                                                                         \
        L.pub[2 * S.pub[r]] = YR[r];                                                               \
        L.pub[2 * S.pub[r] + 1] = YI[r];                                                           \
      }                                                                                            \
      if (s < P.S) {                                                                               \
        L.kap[2 * s] = KR[r];                                                                      \
        L.kap[2 * s + 1] = KI[r];                                                                  \
      }                                                                                            \
    }                                                                                              \
    __syncthreads();                                                                               \
    double tci = 0.0;                                                                              \
    adjoint_stage<R>(P, L, S, KR, KI, YR, YI, L.coef + (i) * KC, mu_r[i], mu_i[i], g, tci, tid, W); \
    tbl += tci;                                                                                    \
    hbl += T::c(i) * tci;                                                                          \
    __syncthreads();                                                                               \
  }
          double lpr[R], lpi[R];
#pragma unroll
          for (int r = 0; r < R; ++r) { lpr[r] = lr[r]; lpi[r] = li[r]; }
          if (errb != 0.0) {
            // stage 7: k7 = f(x + h, yn) enters the error estimate only
            double kr7[R], ki7[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
              kr7[r] = h * (T::e(7) * nur[r]);
              ki7[r] = S.diag(r) ? 0.0 : h * (T::e(7) * nui[r]);
            }
            LB_ADJ_REV(7, kr7, ki7, yn_r, yn_i)
#pragma unroll
            for (int r = 0; r < R; ++r) {
              lpr[r] += mu_r[7][r];
              lpi[r] = S.diag(r) ? 0.0 : lpi[r] + mu_i[7][r];
            }
          }
#define LB_ADJ_STAGE(i)                                                                             \
  {                                                                                                 \
    double kr[R], ki[R], yr[R], yi[R];                                                              \
    _Pragma("unroll") for (int r = 0; r < R; ++r) {                                                 \
      double ar = T::a(7, i) * lpr[r] + T::e(i) * nur[r], ai = T::a(7, i) * lpi[r] + T::e(i) * nui[r]; \
      _Pragma("unroll") for (int j = i + 1; j <= 6; ++j) {                                          \
        if (T::a(j, i) != 0.0) {                                                                    \
          ar = fma(T::a(j, i), mu_r[j][r], ar);                                                     \
          ai = fma(T::a(j, i), mu_i[j][r], ai);                                                     \
        }                                                                                           \
      }                                                                                             \
      kr[r] = h * ar;                                                                               \
      ki[r] = S.diag(r) ? 0.0 : h * ai;                                                             \
    }                                                                                               \
    if (i == 1) {                                                                                   \
      _Pragma("unroll") for (int r = 0; r < R; ++r) { yr[r] = y_r[r]; yi[r] = y_i[r]; }             \
    } else {                                                                                        \
      stage_combo<R, SOLVER, (i > 1 ? i : 2)>(h, y_r, y_i, k_r, k_i, yr, yi);                        \
    }                                                                                               \
    LB_ADJ_REV(i, kr, ki, yr, yi)                                                                   \
  }
          LB_ADJ_STAGE(6)
          LB_ADJ_STAGE(5)
          LB_ADJ_STAGE(4)
          LB_ADJ_STAGE(3)
          LB_ADJ_STAGE(2)
          LB_ADJ_STAGE(1)
#undef LB_ADJ_STAGE
#undef LB_ADJ_REV
#pragma unroll
          for (int r = 0; r < R; ++r) {
            double sr = lpr[r], si = lpi[r];
#pragma unroll
            for (int i = 1; i <= 6; ++i) {
              sr += mu_r[i][r];
              si += mu_i[i][r];
            }
            ybr[r] += sr;
            ybi[r] += si;
          }
          if (adaptive) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              double kk_r[8][1], kk_i[8][1], mm_r[8], mm_i[8];
#pragma unroll
              for (int j = 0; j < 8; ++j) {
                kk_r[j][0] = k_r[j][r];
                kk_i[j][0] = k_i[j][r];
                mm_r[j] = mu_r[j][r];
                mm_i[j] = mu_i[j][r];
              }
              hbl += h_products<SOLVER>(kk_r, kk_i, lpr[r], lpi[r], nur[r], nui[r], mm_r, mm_i);
            }
            hb += group_sum(hbl, L.red, tid, W);
            Tb += group_sum(tbl, L.red, tid, W);
            if (fl & LB_TAPE_CLIPPED) {
              Tb -= hb;
              Hb = 0.0;
            } else {
              Hb = hb;
            }
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            lam_r[r] = acc ? ybr[r] : lam_r[r] + ybr[r];
            lam_i[r] = S.diag(r) ? 0.0 : (acc ? ybi[r] : lam_i[r] + ybi[r]);
          }
          if (acc) Qb = qnew;
        }
      }
    }
    // ---- dI/dp: one block reduction per parameter
#pragma unroll
    for (int m = 0; m < LB_MAX_DIFF; ++m) {
      if (m >= np) break;
      const double v = wave_sum(g[m]);
      if ((tid & 63) == 0) L.red[m * 16 + (tid >> 6)] = v;
    }
    __syncthreads();
    if (tid == 0) {
      const int nw = W >> 6;
      for (int m = 0; m < np; ++m) {
        double tot = 0.0;
        for (int b = 0; b < nw; ++b) tot += L.red[m * 16 + b];
        C.out_grad[traj * np + m] = ok ? tot : __builtin_nan("");
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint for plans of one 64-slot block (lane = slot), several waves per
// trajectory.
//
// The same adjoint as lb_adjoint_kernel, arranged for latency: the gather and
// transposed term tables live in LDS for the whole launch.  One wave prepares
// checkpoint segments: it evaluates the coefficient rows of a segment (every
// stage time of its attempts is on the tape) in one pass, one (attempt,
// stage) per lane, each lane with a private copy of the parameter slots and
// its own expression stack, and recomputes the segment with the
// register/cross-lane RHS of the single-wave solver (the same arithmetic, so
// the stage derivatives are those of the recorded solve).  LB_AW_SWEEPERS
// waves sweep the attempts with the same per-element state; each applies its
// share of the transposed rows of every reverse stage (kappa of other
// elements by cross-lane permutes, each term contracted with its
// coefficient's parameter and time derivatives), and the partial sums of mu
// and of the step-size and time cotangents meet in LDS in a fixed order, so
// every sweeping wave continues with the same values.  The self terms of an
// element (time-independent coefficients) and their parameter derivatives
// are folded once per trajectory and counted by the first wave.
// ---------------------------------------------------------------------------
namespace {

struct LdsAW {
  double2* gf;     // g_rows x 64 gather factors
  double2* tf;     // t_rows x 64 transposed factors
  dcpx* coef0;     // KC: coefficient row at t0 (time-independent entries)
  dcpx* rows;      // nbuf x every * 7 rows of stride KCP: coefficient rows of a segment
  dcpx* slots0;    // n_slots: parameter slots at t0
  dcpx* lslots;    // 64 x SSP: per-lane parameter slots
  dcpx* lvm;       // 64 x VSP: per-lane expression stacks
  double2* selfc;  // 64 folded self coefficients
  dcpx* selfd;     // LB_MAX_DIFF x 64 parameter derivatives of the folded self coefficients
  double* wts;     // n integral weights
  uint32_t* gw;    // g_rows x 64 gather words
  uint32_t* tw;    // t_rows x 64 transposed words
  double2* part;   // 2 x LB_AW_SWEEPERS x 64: per-wave partial mu of a reverse stage (two buffers)
  double* xs;      // 4 * LB_AW_SWEEPERS: per-wave partial step/time cotangents (two buffers),
                   // then LB_AW_SWEEPERS * LB_MAX_DIFF per-wave gradient sums
  int* misc;
};

// Odd strides (in 16-byte units) spread the lanes' rows over the LDS banks.
__device__ __forceinline__ int odd_up(int v) { return v | 1; }

__device__ __forceinline__ LdsAW carve_aw(const DevPlan& P, int every, int nbuf) {
  extern __shared__ double smem[];
  const int g = cst(P.blk_rows)[0], t = cst(P.t_blk_rows)[0], KC = P.KD + 1;
  char* q = reinterpret_cast<char*>(smem);
  LdsAW L;
  L.gf = reinterpret_cast<double2*>(q); q += (size_t)g * 64 * 16;
  L.tf = reinterpret_cast<double2*>(q); q += (size_t)t * 64 * 16;
  L.coef0 = reinterpret_cast<dcpx*>(q); q += (size_t)KC * 16;
  L.rows = reinterpret_cast<dcpx*>(q); q += (size_t)nbuf * every * 7 * odd_up(KC) * 16;
  L.slots0 = reinterpret_cast<dcpx*>(q); q += (size_t)P.n_slots * 16;
  L.lslots = reinterpret_cast<dcpx*>(q); q += (size_t)64 * odd_up(P.n_slots) * 16;
  L.lvm = reinterpret_cast<dcpx*>(q); q += (size_t)64 * odd_up(P.vm_depth) * 16;
  L.selfc = reinterpret_cast<double2*>(q); q += (size_t)64 * 16;
  L.selfd = reinterpret_cast<dcpx*>(q); q += (size_t)LB_MAX_DIFF * 64 * 16;
  L.wts = reinterpret_cast<double*>(q); q += (size_t)((P.n + 1) & ~1) * 8;
  L.gw = reinterpret_cast<uint32_t*>(q); q += (size_t)g * 64 * 4;
  L.tw = reinterpret_cast<uint32_t*>(q); q += (size_t)t * 64 * 4;
  L.part = reinterpret_cast<double2*>(q); q += (size_t)2 * LB_AW_SWEEPERS * 64 * 16;
  L.xs = reinterpret_cast<double*>(q); q += (size_t)(4 + LB_MAX_DIFF) * LB_AW_SWEEPERS * 8;
  L.misc = reinterpret_cast<int*>(q);
  return L;
}

// One reverse stage over the transposed rows [r0, r1) of this lane's element:
// its part of mu = A(t_i)^T kappa, of g_m += Re(conj(kappa)^T (dA(t_i)/dp_m)
// Y_i) and (a plan with a time column) of tc = Re(conj(kappa)^T (dA(t_i)/dt)
// Y_i); `own` adds the self terms.  `cf` is the stage's coefficient row
// (values, the constant 1 at K, then d c_k / d p_m at K + 1 + k * nd + m, the
// time column last); each term is contracted with its coefficient's
// derivatives as it is read.
constexpr int kAwChunk = 4;  // transposed rows fetched together

__device__ __forceinline__ void adjoint_wave_stage(const DevPlan& P, const LdsAW& L, int r0, int r1, bool own,
                                                   int lane, double kr, double ki, double yr, double yi, bool diag,
                                                   double2 sv, const dcpx* cf, double& mr, double& mi,
                                                   double (&g)[LB_MAX_DIFF], double& tc) {
  const int nd = P.n_diff, np = P.n_diff - P.diff_time;
  const dcpx* dc = cf + P.K + 1;
  const double ckr = kr, cki = -ki;  // conj(kappa)
  mr = 0.0;
  mi = 0.0;
  if (own) {
    mr = ckr * sv.x - cki * sv.y;
    mi = diag ? 0.0 : -(ckr * sv.y + cki * sv.x);
    const dcpx cx = cmul(mk(ckr, cki), mk(yr, diag ? 0.0 : yi));
#pragma unroll
    for (int m = 0; m < LB_MAX_DIFF; ++m) {
      if (m >= np) break;
      const dcpx sd = L.selfd[m * 64 + lane];
      g[m] += cx.re * sd.re - cx.im * sd.im;
    }
  }
  constexpr int CH = kAwChunk;
  for (int c0 = r0; c0 < r1; c0 += CH) {
    uint32_t w[CH];
    double2 f[CH];
    double kor[CH], koi[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (c0 + j < r1) {
        w[j] = L.tw[(c0 + j) * 64 + lane];
        f[j] = L.tf[(c0 + j) * 64 + lane];
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (c0 + j < r1) {
        const int o = (int)(w[j] & 63u);
        kor[j] = __shfl(kr, o, 64);
        koi[j] = __shfl(ki, o, 64);
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (c0 + j >= r1) continue;
      const uint32_t wj = w[j], cid = wj >> 24;
      const bool stat = cid == LB_STATIC_COEF, cj = (wj & (1u << 22)) != 0;
      const uint32_t k = stat ? 0u : cid;
      // conj(kappa_o) f, then times the coefficient (1 for a static term)
      const dcpx fo = cmul(mk(kor[j], -koi[j]), mk(f[j].x, f[j].y));
      dcpx c = cf[k];
      c.re = stat ? 1.0 : c.re;
      c.im = stat ? 0.0 : (cj ? -c.im : c.im);
      const dcpx hh = cmul(fo, c);
      mr += hh.re;
      mi += (wj & (1u << 20)) ? 0.0 : ((wj & (1u << 21)) ? hh.im : -hh.im);
      // the term's contraction with its coefficient's derivatives, read from
      // the stage row (weight zero for a static term); Re(qx d') with
      // d' = conj(d) for a conjugated coefficient is Re(conj(qx) d)
      double xi = (wj & (1u << 20)) ? 0.0 : yi;
      xi = (wj & (1u << 21)) ? -xi : xi;
      const double zw = stat ? 0.0 : 1.0;
      const dcpx qx = cmul(fo, mk(yr * zw, xi * zw));
      const double qi = cj ? -qx.im : qx.im;
      const dcpx* dk = dc + k * nd;
#pragma unroll
      for (int m = 0; m < LB_MAX_DIFF; ++m) {
        if (m >= np) break;
        const dcpx d = dk[m];
        g[m] += qx.re * d.re - qi * d.im;
      }
      if (P.diff_time) {
        const dcpx d = dk[np];
        tc += qx.re * d.re - qi * d.im;
      }
    }
  }
}

}  // namespace

template <int SOLVER, int G, int WAVES>
__global__ __launch_bounds__(64 * (LB_AW_SWEEPERS + 1), WAVES) void lb_adjoint_wave_kernel(const DevPlan P,
                                                                                          const DevCall C) {
  using T = Tab<SOLVER>;
  const int tid = threadIdx.x, lane = tid & 63;
  // wave index (uniform) and waves per trajectory: with LB_AW_SWEEPERS + 1,
  // waves 0 .. LB_AW_SWEEPERS - 1 sweep a segment's attempts while the last
  // evaluates the coefficient rows of the next segment and recomputes it into
  // the other of two buffers; with LB_AW_SWEEPERS (or fewer), wave 0 does
  // that work before the sweep, into one buffer
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = (int)(blockDim.x >> 6);
  const bool ahead = nw > LB_AW_SWEEPERS;
  const int nrev = ahead ? LB_AW_SWEEPERS : nw, prod = ahead ? LB_AW_SWEEPERS : 0, nbuf = ahead ? 2 : 1;
  const bool rev = wv < nrev, lead = wv == 0;
  const int every = C.ckpt_every;
  const LdsAW L = carve_aw(P, every, nbuf);
  const int KC = P.KD + 1, KCP = odd_up(KC), nd = P.n_diff, np = P.n_diff - P.diff_time;
  const bool adaptive = C.adj_mode == LB_ADJOINT_ADAPTIVE;
  const double dimf = (double)P.dim;
  const int grows = cst(P.blk_rows)[0], trows = cst(P.t_blk_rows)[0];
  // this wave's transposed rows: the chunks of kAwChunk rows in equal runs,
  // the first sweeping wave's first
  const int nch = (trows + kAwChunk - 1) / kAwChunk, cpw = (nch + nrev - 1) / nrev;
  const int tr0 = min(trows, wv * cpw * kAwChunk), tr1 = min(trows, (wv + 1) * cpw * kAwChunk);
  Slots<1> S;
  load_slots<1>(P, S, lane, 64);
  const bool diag = S.diag(0), valid = S.valid(0);
  for (int i = tid; i < grows * 64; i += (int)blockDim.x) {
    L.gf[i] = reinterpret_cast<const double2*>(P.g_f)[i];
    L.gw[i] = P.g_word[i];
  }
  for (int i = tid; i < trows * 64; i += (int)blockDim.x) {
    L.tf[i] = reinterpret_cast<const double2*>(P.t_f)[i];
    L.tw[i] = P.t_word[i];
  }
  // misc[1 + w]: exchanges posted by sweeping wave w; misc[LB_AW_SWEEPERS + 1]:
  // 1 when an exchange wait ran out (the gradients become NaN)
  int* const posts = L.misc + 1;
  int* const broke = L.misc + 1 + LB_AW_SWEEPERS;
  if (tid == 0) {
    for (int w = 0; w < LB_AW_SWEEPERS; ++w) posts[w] = 0;
    *broke = C.adj_wait_limit < 0 ? 1 : 0;
  }
  const int wait_limit = max(C.adj_wait_limit, 0);
  int par = 0, posted = 0;  // exchange buffer and count of this wave's exchanges
  bool broken = false;
  // mu (or a scalar pair) summed over the sweeping waves in a fixed tree,
  // ((w0 + w1) + (w2 + w3)).  The waves meet through LDS flags, not a
  // workgroup barrier (the preparing wave works on the next segment
  // meanwhile): each posts its part and its exchange count, then waits until
  // every other sweeping wave has posted the same count.  Two buffers
  // alternate, so a wave rewrites one only after the others have passed the
  // following exchange, that is after they have read it.  The wait is
  // bounded: a wave that waits too long marks the launch broken and every
  // later wait returns at once, so the kernel always ends.
  auto meet = [&](double a, double b, double& sa, double& sb, bool per_lane) {
    if (nrev == 1) {
      sa = a;
      sb = b;
      return;
    }
    double2* const pp = L.part + (size_t)par * LB_AW_SWEEPERS * 64;
    double* const xp = L.xs + par * 2 * LB_AW_SWEEPERS;
    if (per_lane) {
      pp[wv * 64 + lane] = double2{a, b};
    } else if (lane == 0) {
      xp[wv * 2] = a;
      xp[wv * 2 + 1] = b;
    }
    ++posted;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(&posts[wv], posted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (int w = 0; w < nrev; ++w) {
      if (w == wv) continue;
      for (int it = 0; !broken; ++it) {
        if (__hip_atomic_load(&posts[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= posted) break;
        if (__hip_atomic_load(broke, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0 || it >= wait_limit) {
          broken = true;
          if (lane == 0) __hip_atomic_store(broke, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double ra = 0.0, rb = 0.0;
    if (per_lane) {
      const double2 u0 = pp[lane], u1 = pp[64 + lane];
      ra = u0.x + u1.x;
      rb = u0.y + u1.y;
      if (nrev > 2) {
        const double2 u2 = pp[128 + lane], u3 = nrev > 3 ? pp[192 + lane] : double2{0.0, 0.0};
        ra = ra + (u2.x + u3.x);
        rb = rb + (u2.y + u3.y);
      }
    } else {
      ra = xp[0] + xp[2];
      rb = xp[1] + xp[3];
      if (nrev > 2) {
        ra = ra + (xp[4] + (nrev > 3 ? xp[6] : 0.0));
        rb = rb + (xp[5] + (nrev > 3 ? xp[7] : 0.0));
      }
    }
    sa = ra;
    sb = rb;
    par ^= 1;
  };
  Lds LV{};  // the gather tables as the single-wave RHS reads them
  LV.gw = L.gw;
  LV.gf = L.gf;
  dcpx* const myslots = L.lslots + lane * odd_up(P.n_slots);
  dcpx* const myvm = L.lvm + lane * odd_up(P.vm_depth);
  // workspace per attempt: y and k1..k7 of this lane's element, per buffer
  double2* const segW = reinterpret_cast<double2*>(C.seg_y) + (size_t)blockIdx.x * nbuf * every * 8 * 64;
  const int64_t ncap = (C.tape_cap + every - 1) / every;

  for (;;) {
    __syncthreads();
    if (tid == 0) L.misc[0] = atomicAdd(C.queue, 1);
    __syncthreads();
    const int64_t local = L.misc[0];
    if (local >= C.n_traj) break;
    // queue position -> trajectory (longest-first orders); entries outside
    // the batch are skipped, so a bad order cannot address past the arrays
    const int64_t traj = C.queue_order ? (int64_t)C.queue_order[local + C.traj_offset] : local + C.traj_offset;
    if (traj < 0 || traj >= C.n_total) continue;

    // ---- parameters and coefficients at t0, weights (the first wave)
    if (lead)
      for (int i = lane; i < P.n_slots; i += 64)
        L.slots0[i] = i < P.n_base ? mk(P.base[2 * i], P.base[2 * i + 1]) : mk(0.0, 0.0);
    __syncthreads();
    if (tid == 0) {
      apply_overrides(C, traj, L.slots0);
      eval_coefficients(P, L.slots0, L.coef0, C.t0, false, myvm, P.KD);
    }
    if (lead)
      for (int i = lane; i < P.n; i += 64) {
        double w = 0.0;
        for (int q = 0; q < C.n_weights; ++q)
          if (C.weight_index[q] == i) w += C.weight_value[q];
        L.wts[i] = w;
      }
    __syncthreads();
    if (wv == prod)
      for (int i = 0; i < P.n_slots; ++i) myslots[i] = L.slots0[i];
    if (lead) fold_self(P, lane, 64, L.coef0, reinterpret_cast<double*>(L.selfc));
    // parameter derivatives of the folded self coefficient
    for (int m = 0; lead && m < np; ++m) {
      dcpx acc = mk(0.0, 0.0);
      for (int j = 0; j < P.max_self; ++j) {
        const size_t e = (size_t)lane * P.max_self + j;
        const int id = P.self_coef[e];
        if (id == -2) break;
        if (id < 0) continue;
        dcpx d = L.coef0[P.K + 1 + id * nd + m];
        if (P.self_conj[e]) d.im = -d.im;
        acc = cadd(acc, cmul(mk(P.self_f[2 * e], P.self_f[2 * e + 1]), d));
      }
      L.selfd[m * 64 + lane] = acc;
    }
    __syncthreads();
    const double2 selfv = L.selfc[lane];

    double g[LB_MAX_DIFF];
#pragma unroll
    for (int m = 0; m < LB_MAX_DIFF; ++m) g[m] = 0.0;
    const int N = C.tape_n[traj];
    const bool ok = N > 0 && N <= C.tape_cap;
    if (ok) {
      const double* tt = C.tape_t + traj * C.tape_cap;
      const double* th = C.tape_h + traj * C.tape_cap;
      const double* tcr = C.tape_c + traj * C.tape_cap * 3;
      const int32_t* tfl = C.tape_f + traj * C.tape_cap;
      const double wme = (valid && diag) ? L.wts[S.state(0)] : 0.0;
      // cotangents: lam of the state at the start of the current step; Tb of
      // its start time; Hb of the step size proposed into the attempt after
      // the current one (before its clip); Qb of fac_old after the current step
      double lam_r = 0.0, lam_i = 0.0, Tb = 0.0, Hb = 0.0, Qb = 0.0;
      const int nseg = (N + every - 1) / every;
      // Phase ph produces segment nseg - 1 - ph and sweeps the one produced
      // before it (two buffers), or produces and sweeps the same segment (one).
      const int lag = nbuf - 1;
      for (int ph = 0; ph < nseg + lag; ++ph) {
        const int sp = nseg - 1 - ph, sr = sp + lag;
        if (wv == prod && sp >= 0) {
          const int a0 = sp * every, a1 = min(a0 + every, N), ntask = 7 * (a1 - a0);
          dcpx* const rows = L.rows + (size_t)(sp % nbuf) * every * 7 * KCP;
          double2* const wseg = segW + (size_t)(sp % nbuf) * every * 8 * 64;
          // ---- coefficient rows of the segment: task q = attempt (q / 7), stage (q % 7 + 1)
          for (int q = lane; q < ntask; q += 64) {
            const int a = a0 + q / 7, st = q % 7 + 1;
            const double ts = st == 1 ? tt[a] : tt[a] + th[a] * T::c(st);
            dcpx* row = rows + q * KCP;
            for (int j = 0; j < KC; ++j) row[j] = L.coef0[j];
            if (P.time_dependent) eval_coefficients(P, myslots, row, ts, true, myvm, P.KD);
          }
          wave_sync();
          // ---- recompute the segment from its checkpoint, keeping y and k1..k7
          double y_r[1], y_i[1];
          {
            const double2 ck = reinterpret_cast<const double2*>(C.tape_y + ((traj * ncap + sp) * P.S) * 2)[lane];
            y_r[0] = ck.x;
            y_i[0] = ck.y;
          }
          for (int a = a0; a < a1; ++a) {
            const double h = th[a];
            const dcpx* rw = rows + (a - a0) * 7 * KCP;
            double2* wk = wseg + (size_t)(a - a0) * 8 * 64 + lane;
            double k_r[8][1], k_i[8][1], yt_r[1], yt_i[1];
            wk[0] = double2{y_r[0], y_i[0]};
            rhs_wave<G>(LV, grows, lane, selfv, diag, y_r[0], y_i[0], k_r[1][0], k_i[1][0], rw);
            wk[64] = double2{k_r[1][0], k_i[1][0]};
#define LB_AW_FWD(st)                                                                           \
  stage_combo<1, SOLVER, st>(h, y_r, y_i, k_r, k_i, yt_r, yt_i);                                \
  rhs_wave<G>(LV, grows, lane, selfv, diag, yt_r[0], yt_i[0], k_r[st][0], k_i[st][0], rw + (st - 1) * KCP); \
  wk[(st) * 64] = double2{k_r[st][0], k_i[st][0]};
            LB_AW_FWD(2)
            LB_AW_FWD(3)
            LB_AW_FWD(4)
            LB_AW_FWD(5)
            LB_AW_FWD(6)
            LB_AW_FWD(7)
#undef LB_AW_FWD
            if (tfl[a] & LB_TAPE_ACCEPTED) {
              // stage_combo<7> left the candidate solution in yt
              y_r[0] = yt_r[0];
              y_i[0] = yt_i[0];
            }
          }
        }
        if (lag == 0) __syncthreads();  // one buffer: the sweep reads what was just written
        if (!rev || sr < 0 || sr >= nseg) {
          __syncthreads();
          continue;
        }
        const int a0 = sr * every, a1 = min(a0 + every, N);
        const dcpx* const rows = L.rows + (size_t)(sr % nbuf) * every * 7 * KCP;
        const double2* const wseg = segW + (size_t)(sr % nbuf) * every * 8 * 64;
        // ---- reverse sweep over the segment's attempts, last first (both
        // sweeping waves, the same control flow: every branch below depends
        // on the tape and on cotangents both waves hold)
        for (int a = a1 - 1; a >= a0; --a) {
          const int fl = tfl[a];
          const bool acc = (fl & LB_TAPE_ACCEPTED) != 0;
          if (!adaptive && !acc) continue;
          const double h = th[a];
          double hb = 0.0, errb = 0.0, qnew = Qb;
          if (adaptive) {
            double dh, de, dq;
            ctrl_partials(fl, h, tcr[3 * a], tcr[3 * a + 1], tcr[3 * a + 2], dh, de, dq);
            if (acc) {
              errb = Hb * de + ((fl & LB_TAPE_ERRQ) ? Qb : 0.0);
              hb = Hb * dh + Tb;
              qnew = Hb * dq;
            } else {
              errb = Hb * de;
              hb = Hb * dh;
            }
            // a rejected attempt whose proposal nothing downstream depends on
            if (!acc && errb == 0.0 && hb == 0.0) {
              if (fl & LB_TAPE_CLIPPED) Hb = 0.0;
              continue;
            }
          }
          const dcpx* rw = rows + (a - a0) * 7 * KCP;
          const double2* wk = wseg + (size_t)(a - a0) * 8 * 64 + lane;
          double y_r[1], y_i[1], k_r[8][1], k_i[8][1], yn_r[1], yn_i[1];
          {
            const double2 v = wk[0];
            y_r[0] = v.x;
            y_i[0] = v.y;
          }
#pragma unroll
          for (int j = 1; j <= 7; ++j) {
            const double2 v = wk[j * 64];
            k_r[j][0] = v.x;
            k_i[j][0] = v.y;
          }
          k_r[0][0] = 0.0;
          k_i[0][0] = 0.0;
          stage_combo<1, SOLVER, 7>(h, y_r, y_i, k_r, k_i, yn_r, yn_i);
          double lr = acc ? lam_r : 0.0, li = acc ? lam_i : 0.0;
          double ybr = 0.0, ybi = 0.0, nur = 0.0, nui = 0.0, hbl = 0.0, tbl = 0.0;
          if (acc && (!(fl & LB_TAPE_FIRST) || C.save_start)) {
            // 0.5 (v_n + v_{n+1}) h with v = w . y
            lr += 0.5 * h * wme;
            ybr += 0.5 * h * wme;
            if (lead) hbl += 0.5 * wme * (y_r[0] + yn_r[0]);
          }
          if (errb != 0.0 && valid) {
            const double err = tcr[3 * a];
            double er = T::e(1) * k_r[1][0], ei = T::e(1) * k_i[1][0];
#pragma unroll
            for (int j = 2; j <= 7; ++j) {
              if (T::e(j) != 0.0) {
                er = fma(T::e(j), k_r[j][0], er);
                ei = fma(T::e(j), k_i[j][0], ei);
              }
            }
            err_component(errb, err, dimf, C.abstol, C.reltol, er * h, y_r[0], yn_r[0], nur, ybr, lr);
            if (!diag) err_component(errb, err, dimf, C.abstol, C.reltol, ei * h, y_i[0], yn_i[0], nui, ybi, li);
          }
          double mu_r[8], mu_i[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) { mu_r[i] = 0.0; mu_i[i] = 0.0; }
          double lpr = lr, lpi = li;
          if (errb != 0.0) {
            // stage 7: k7 = f(x + h, yn) enters the error estimate only
            double tc7 = 0.0, pr, pi;
            adjoint_wave_stage(P, L, tr0, tr1, lead, lane, h * (T::e(7) * nur), diag ? 0.0 : h * (T::e(7) * nui),
                               yn_r[0], yn_i[0], diag, selfv, rw + 6 * KCP, pr, pi, g, tc7);
            meet(pr, pi, mu_r[7], mu_i[7], true);
            tbl += tc7;
            hbl += tc7;
            lpr += mu_r[7];
            lpi = diag ? 0.0 : lpi + mu_i[7];
          }
#define LB_AW_STAGE(i)                                                                                  \
  {                                                                                                     \
    double ar = T::a(7, i) * lpr + T::e(i) * nur, ai = T::a(7, i) * lpi + T::e(i) * nui;                \
    _Pragma("unroll") for (int j = i + 1; j <= 6; ++j) {                                                \
      if (T::a(j, i) != 0.0) {                                                                          \
        ar = fma(T::a(j, i), mu_r[j], ar);                                                              \
        ai = fma(T::a(j, i), mu_i[j], ai);                                                              \
      }                                                                                                 \
    }                                                                                                   \
    double ys_r[1], ys_i[1];                                                                            \
    if (i == 1) {                                                                                       \
      ys_r[0] = y_r[0];                                                                                 \
      ys_i[0] = y_i[0];                                                                                 \
    } else {                                                                                            \
      stage_combo<1, SOLVER, (i > 1 ? i : 2)>(h, y_r, y_i, k_r, k_i, ys_r, ys_i);                        \
    }                                                                                                   \
    double tci = 0.0, pr, pi;                                                                           \
    adjoint_wave_stage(P, L, tr0, tr1, lead, lane, h * ar, diag ? 0.0 : h * ai, ys_r[0], ys_i[0], diag, \
                       selfv, rw + (i - 1) * KCP, pr, pi, g, tci);                                      \
    meet(pr, pi, mu_r[i], mu_i[i], true);                                                               \
    tbl += tci;                                                                                         \
    hbl += T::c(i) * tci;                                                                               \
  }
          LB_AW_STAGE(6)
          LB_AW_STAGE(5)
          LB_AW_STAGE(4)
          LB_AW_STAGE(3)
          LB_AW_STAGE(2)
          LB_AW_STAGE(1)
#undef LB_AW_STAGE
          double sr = lpr, si = lpi;
#pragma unroll
          for (int i = 1; i <= 6; ++i) {
            sr += mu_r[i];
            si += mu_i[i];
          }
          ybr += sr;
          ybi += si;
          if (adaptive) {
            if (lead) hbl += h_products<SOLVER>(k_r, k_i, lpr, lpi, nur, nui, mu_r, mu_i);
            double sh, st;
            meet(wave_sum(hbl), wave_sum(tbl), sh, st, false);
            hb += sh;
            Tb += st;
            if (fl & LB_TAPE_CLIPPED) {
              Tb -= hb;
              Hb = 0.0;
            } else {
              Hb = hb;
            }
          }
          lam_r = acc ? ybr : lam_r + ybr;
          lam_i = diag ? 0.0 : (acc ? ybi : lam_i + ybi);
          if (acc) Qb = qnew;
        }
        __syncthreads();
      }
    }
    // ---- dI/dp: one wave sum per parameter and sweeping wave, the waves in order
#pragma unroll
    for (int m = 0; m < LB_MAX_DIFF; ++m) {
      if (m >= np) break;
      const double v = wave_sum(g[m]);
      if (rev && lane == 0) L.xs[4 * LB_AW_SWEEPERS + wv * LB_MAX_DIFF + m] = v;
    }
    __syncthreads();
    if (tid == 0) {
      const bool fine = ok && *broke == 0;
      const double* gs = L.xs + 4 * LB_AW_SWEEPERS;
      for (int m = 0; m < np; ++m) {
        double v = gs[m];
        for (int w = 1; w < nrev; ++w) v += gs[w * LB_MAX_DIFF + m];
        C.out_grad[traj * np + m] = fine ? v : __builtin_nan("");
      }
    }
  }
}


// ---------------------------------------------------------------------------
// Batched RHS (one evaluation per workgroup).
// ---------------------------------------------------------------------------
template <int R, int MAXW, int WAVES>
__global__ __launch_bounds__(MAXW, WAVES) void lb_rhs_kernel(const DevPlan P, const DevRhsCall C) {
  const int tid = threadIdx.x, W = blockDim.x;
  const Lds L = carve(P);
  Slots<R> S;
  load_slots<R>(P, S, tid, W);
  stage_tables(P, L, tid, W);
  for (int64_t b = blockIdx.x; b < C.batch; b += gridDim.x) {
    __syncthreads();
    init_slots(P, L.slots, tid, W);
    if (tid == 0) {
      for (int j = 0; j < C.n_override; ++j)
        L.slots[C.override_slot[j]] = mk(C.override_values[2 * (b * C.n_override + j)],
                                         C.override_values[2 * (b * C.n_override + j) + 1]);
      eval_coefficients(P, L.slots, L.coef, C.t[b], false, L.vm);
    }
    __syncthreads();
    fold_self(P, tid, W, L.coef, L.selfc);
    double xr[R], xi[R], kr[R], ki[R];
    const double* y = C.y + b * (int64_t)P.dim;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int s = tid + W * r;
      xr[r] = S.valid(r) ? y[P.slot_re[s]] : 0.0;
      xi[r] = (S.valid(r) && !S.diag(r)) ? y[P.slot_im[s]] : 0.0;
    }
    rhs_eval<R>(P, L, S, xr, xi, kr, ki, L.pub, L.coef, tid, W);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!S.valid(r)) continue;
      const int s = tid + W * r;
      C.out[b * (int64_t)P.dim + P.slot_re[s]] = kr[r];
      if (!S.diag(r)) C.out[b * (int64_t)P.dim + P.slot_im[s]] = ki[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Effective-Hamiltonian Lindblad path (rust/src/effective_lindblad/).  The
// state is real, y = [Re rho.flat, Im rho.flat] (D = 2 n^2), and
//   dy = A_comb(x) y + (Omega / 2) A_opt(x) y + delta A_det(x) y
// (rhs.rs:86-138): three CSR operators whose stored entries are cubics in
// dx = x - grid[k] on grid interval k, with x, Omega, delta the parameter
// slots at t.  One workgroup per trajectory from the work queue, R rows per
// lane, the stage input published to LDS (double-buffered) for the column
// gathers.  The parameter graph runs in six lanes per step, one per stage
// time; lane 0 then finds the grid intervals in the order the reference's
// RHS calls would, carrying its cached interval hint (rhs.rs:47-84).  The
// integrator is the expanded-sparse one: Dopri5 / Tsit5 tableaus,
// controller, stiffness test, save plan and dense output.
// ---------------------------------------------------------------------------
#if LB_PART == 1
namespace {

struct LdsEff {
  double* ybuf;  // 2 x S stage inputs
  dcpx* slots;   // 8 x n_slots: t0 / hinit copy and the six stage times
  dcpx* vm;      // 8 x vm_depth
  double* red;   // 4 x nblk block sums
  double* sp;    // 8 x 4 per RHS site: Omega/2, delta, dx, interval; [4 s + 2] field value before lookup
  double* ev;    // [0] field value of a failed lookup
  int* misc;     // [0] queue value, [1] status, [2] interval hint
};

// Must match lds_bytes_eff() in lb_host.cpp.
__device__ __forceinline__ LdsEff carve_eff(const DevEffPlan& P) {
  extern __shared__ double smem[];
  LdsEff L;
  L.ybuf = smem;
  L.slots = reinterpret_cast<dcpx*>(L.ybuf + 2 * (size_t)P.S);
  L.vm = L.slots + 8 * (size_t)P.n_slots;
  L.red = reinterpret_cast<double*>(L.vm + 8 * (size_t)P.vm_depth);
  L.sp = L.red + 4 * (size_t)P.nblk;
  L.ev = L.sp + 32;
  L.misc = reinterpret_cast<int*>(L.ev + 2);
  return L;
}

__device__ __forceinline__ void eff_block_store(const DevEffPlan& P, int tid, int W, int r, double* red, int q,
                                                double value) {
  const int s = tid + W * r;
  const double sum = wave_sum(value);
  if ((tid & 63) == 0 && s < P.S) red[q * P.nblk + (s >> 6)] = sum;
}

__device__ __forceinline__ double eff_block_total(const DevEffPlan& P, const double* red, int q) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
  for (int b = lane; b < P.nblk; b += 64) v += red[q * P.nblk + b];
  return wave_sum(v);
}

// Parameter graph at t with the trajectory's (real) overrides, into slot copy
// `sl` (evaluate_with_overrides_into); returns the field value and stores
// Omega/2 and delta in sp[0], sp[1].  One lane.
__device__ __forceinline__ double eff_params(const DevEffPlan& P, const DevEffCall& C, int64_t traj, dcpx* sl,
                                             double t, dcpx* stk, double* sp) {
  for (int i = 0; i < P.n_slots; ++i) sl[i] = i < P.n_base ? mk(P.base[2 * i], P.base[2 * i + 1]) : mk(0.0, 0.0);
  for (int j = 0; j < C.n_override; ++j)
    sl[C.override_slot[j]] = mk(C.override_values ? C.override_values[traj * C.n_override + j] : 0.0, 0.0);
  const __attribute__((address_space(4))) DevExpr* ex = cst(P.comp);
  for (int c = 0; c < P.n_comp; ++c) {
    const DevExpr r = fetch_expr(ex, c);
    sl[r.dst] = vm_eval(cst(P.instr), r.start, r.end, sl, t, stk);
  }
  sp[0] = 0.5 * sl[P.rabi_slot].re;
  sp[1] = sl[P.det_slot].re;
  return sl[P.field_slot].re;
}

// find_interval_cached (rhs.rs:47-84): the interval of field value x and
// dx = x - grid[interval], with the trajectory's hint; -1 / -2 below / above
// the grid by more than its tolerance.  One lane.
__device__ __forceinline__ int eff_interval(const DevEffPlan& P, double x, int& hint, double& dx) {
  const int n = P.G;
  const double g0 = P.grid[0], gn = P.grid[n - 1];
  const double tol = 1e-10 * fmax(fabs(gn - g0), 1.0);
  if (x < g0 - tol) return -1;
  if (x > gn + tol) return -2;
  const double xc = x < g0 ? g0 : (x > gn ? gn : x);
  int idx;
  if (xc <= g0) {
    hint = 0;
    idx = 0;
  } else if (xc >= gn) {
    hint = n - 2;
    idx = n - 2;
  } else if (hint < n - 1 && P.grid[hint] <= xc && xc <= P.grid[hint + 1]) {
    idx = hint;
  } else {
    int lo = 0, hi = n;  // first grid point above xc
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (P.grid[mid] <= xc) lo = mid + 1;
      else hi = mid;
    }
    idx = lo > 0 ? lo - 1 : 0;
    if (idx > n - 2) idx = n - 2;
    hint = idx;
  }
  dx = x - P.grid[idx];
  return idx;
}

// Interval lookup for RHS site s (field value in sp[4 s + 2]); lane 0, in
// call order.  A failed lookup records the first failure and leaves interval
// 0 so the RHS still reads inside the tables.
__device__ __forceinline__ void eff_lookup(const DevEffPlan& P, const LdsEff& L, int s) {
  const double x = L.sp[4 * s + 2];
  int hint = L.misc[2];
  double dx = 0.0;
  const int idx = eff_interval(P, x, hint, dx);
  L.misc[2] = hint;
  if (idx < 0 && L.misc[1] == 0) {
    L.misc[1] = idx == -1 ? LB_TRAJ_FIELD_BELOW : LB_TRAJ_FIELD_ABOVE;
    L.ev[0] = x;
  }
  L.sp[4 * s + 2] = idx < 0 ? 0.0 : dx;
  L.sp[4 * s + 3] = (double)(idx < 0 ? 0 : idx);
}

// dy = sum_o scale_o * sum_p cubic_p(dx) * y[col_p] row by row, each operator
// summed in CSR order and added in the reference's order (rhs.rs:123-135,
// plan.rs:18-40).  Publishes y to LDS buffer `yb` first (one barrier).
template <int R>
__device__ __forceinline__ void eff_rhs(const DevEffPlan& P, double* yb, const double (&y)[R], double (&k)[R],
                                        const double* sp, int tid, int W) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = tid + W * r;
    if (row < P.D) yb[row] = y[r];
  }
  __syncthreads();
  const double half_rabi = sp[0], det = sp[1], dx = sp[2];
  const int idx = (int)sp[3];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = tid + W * r;
    double acc = 0.0;
    if (row < P.D) {
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        const double2* lo = P.coef_lo[o] + (size_t)idx * P.cells[o];
        const double2* hi = P.coef_hi[o] + (size_t)idx * P.cells[o];
        const int* col = P.col[o];
        const int len = P.row_len[o][row];
        const int64_t c0 = P.slice_ptr[o][row >> 6] + (row & 63);
        double sum = 0.0;
        for (int j = 0; j < len; ++j) {  // the row's entries in CSR order
          const int64_t c = c0 + 64 * (int64_t)j;
          const double2 a = lo[c], b = hi[c];
          const double val = a.x + dx * (a.y + dx * (b.x + dx * b.y));
          sum += val * yb[col[c]];
        }
        acc += (o == 0 ? 1.0 : o == 1 ? half_rabi : det) * sum;
      }
    }
    k[r] = acc;
  }
}

template <int R, int SOLVER, int S>
__device__ __forceinline__ void eff_stage(double h, const double (&y)[R], const double (&k)[8][R], double (&o)[R]) {
  using T = Tab<SOLVER>;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double sr = T::a(S, 1) * k[1][r];
#pragma unroll
    for (int j = 2; j < S; ++j)
      if (T::a(S, j) != 0.0) sr = fma(T::a(S, j), k[j][r], sr);
    o[r] = fma(h, sr, y[r]);
  }
}

template <int R, int SOLVER>
__device__ __forceinline__ void eff_dense(double th, double h, const double (&y)[R], const double (&yn)[R],
                                          const double (&k)[8][R], double (&o)[R]) {
  if constexpr (SOLVER == 0) {
    const double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
    const double th1 = 1.0 - th;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double d = yn[r] - y[r];
      const double b = k[1][r] * h - d;
      const double r3 = -k[7][r] * h + d - b;
      const double r4 = h * (D1 * k[1][r] + D3 * k[3][r] + D4 * k[4][r] + D5 * k[5][r] + D6 * k[6][r] + D7 * k[7][r]);
      o[r] = y[r] + (d + (b + (r3 + r4 * th1) * th) * th1) * th;
    }
  } else {
    const double t2 = th * th, t3 = t2 * th, t4 = t3 * th;
    const double b1 = 1.0 * th + -2.763706197274826 * t2 + 2.9132554618219126 * t3 + -1.0530884977290216 * t4;
    const double b2 = 0.1317 * t2 + -0.2234 * t3 + 0.1017 * t4;
    const double b3 = 3.9302962368947516 * t2 + -5.941033872131505 * t3 + 2.490627285651253 * t4;
    const double b4 = -12.411077166933676 * t2 + 30.33818863028232 * t3 + -16.548102889244902 * t4;
    const double b5 = 37.50931341651104 * t2 + -88.1789048947664 * t3 + 47.37952196281928 * t4;
    const double b6 = -27.896526289197286 * t2 + 65.09189467479366 * t3 + -34.87065786149661 * t4;
    const double b7 = 1.5 * t2 + -4.0 * t3 + 2.5 * t4;
#pragma unroll
    for (int r = 0; r < R; ++r)
      o[r] = y[r] + h * (b1 * k[1][r] + b2 * k[2][r] + b3 * k[3][r] + b4 * k[4][r] + b5 * k[5][r] +
                         b6 * k[6][r] + b7 * k[7][r]);
  }
}

}  // namespace

template <int R, int SOLVER, int MAXW, int WAVES>
__global__ __launch_bounds__(MAXW, WAVES) void lb_eff_kernel(const DevEffPlan P, const DevEffCall C) {
  using T = Tab<SOLVER>;
  const int tid = threadIdx.x, W = blockDim.x;
  const LdsEff L = carve_eff(P);
  for (;;) {
    __syncthreads();
    if (tid == 0) L.misc[0] = atomicAdd(C.queue, 1);
    __syncthreads();
    const int64_t traj = L.misc[0];
    if (traj >= C.n_traj) break;
    const double t0 = C.t0, t1 = C.t1;
    if (tid == 0) {
      L.misc[1] = 0;
      L.misc[2] = 0;
    }
    double y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = tid + W * r;
      y[r] = row < P.D ? C.y0[row] : 0.0;
    }
    int64_t accepted = 0, rejected = 0, rhs_calls = 0, n_pushed = 0, si = 0;
    int status = 0, buf = 0;
    double fail_x = 0.0, integral = 0.0, last_t = __builtin_nan(""), last_v = 0.0;
    const bool step_push = C.n_saveat == 0;
    const bool want_int = C.output_kind == LB_EFF_OUT_INTEGRAL;
    // a pushed point: the weighted sum for the integral (in weight order,
    // output.rs:315-330), or an output row
    auto push = [&](double tp, const double (&v)[R]) {
      if (want_int) {
        double* yb = L.ybuf + buf * P.S;
        buf ^= 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tid + W * r;
          if (row < P.D) yb[row] = v[r];
        }
        __syncthreads();
        if (tid == 0) {
          double w = 0.0;
          for (int q = 0; q < C.n_weights; ++q) w += C.weight_value[q] * yb[C.weight_index[q]];
          if (isfinite(last_t)) integral += 0.5 * (last_v + w) * (tp - last_t);
          last_v = w;
        }
        last_t = tp;
        ++n_pushed;
        return;
      }
      if (n_pushed < C.capacity) {
        const int64_t base = (traj * C.capacity + n_pushed) * C.width;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tid + W * r;
          if (row >= P.D) continue;
          if (C.output_kind == LB_EFF_OUT_FULL) {
            C.out_values[base + row] = v[r];
          } else if (C.output_kind == LB_EFF_OUT_POPULATIONS) {
            if (row < P.n * P.n && row % (P.n + 1) == 0) C.out_values[base + row / (P.n + 1)] = v[r];
          } else {
            for (int q = 0; q < C.n_selected; ++q) {
              if (C.sel_re[q] == row) {
                C.out_values[base + 2 * q] = v[r];
                if (C.sel_im[q] < 0) C.out_values[base + 2 * q + 1] = 0.0;
              }
              if (C.sel_im[q] == row) C.out_values[base + 2 * q + 1] = v[r];
            }
          }
        }
        if (tid == 0 && C.out_times) C.out_times[traj * C.capacity + n_pushed] = tp;
      }
      last_t = tp;
      ++n_pushed;
    };
    // field, Omega/2, delta at t for RHS site s (lane 0), then the lookup
    auto site = [&](int s, double t) {
      if (tid == 0) {
        L.sp[4 * s + 2] = eff_params(P, C, traj, L.slots + s * P.n_slots, t, L.vm, L.sp + 4 * s);
        eff_lookup(P, L, s);
      }
      __syncthreads();
      if (L.misc[1] != 0) {
        status = L.misc[1];
        fail_x = L.ev[0];
      }
    };
#define LB_EFF_RHS(yy, kk, s)                              \
  do {                                                     \
    eff_rhs<R>(P, L.ybuf + buf * P.S, yy, kk, L.sp + 4 * (s), tid, W); \
    buf ^= 1;                                              \
  } while (0)

    double k[8][R];
    if (t1 < t0) {
      status = LB_TRAJ_BAD_SPAN;
      fail_x = t0;
    } else if (t1 == t0) {
      if (C.save_start) push(t0, y);
    } else {
      double x = t0, h = 0.0;
      if (isfinite(C.dt) && C.dt > 0.0) {
        h = fmin(C.dt, t1 - t0);
      } else {
        // hinit (rust/src/ode/common.rs:72-118): f0 at t0, f1 at t0 + h0
        double f0[R], yt[R], f1[R];
        site(0, t0);
        if (status == 0) {
          LB_EFF_RHS(y, f0, 0);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int row = tid + W * r;
            const double sc = C.abstol + fabs(y[r]) * C.reltol;
            const double a0 = y[r] / sc, b0 = f0[r] / sc;
            eff_block_store(P, tid, W, r, L.red, 0, row < P.D ? a0 * a0 : 0.0);
            eff_block_store(P, tid, W, r, L.red, 1, row < P.D ? b0 * b0 : 0.0);
          }
          __syncthreads();
          const double d0 = eff_block_total(P, L.red, 0), d1 = eff_block_total(P, L.red, 1), hmax = fabs(t1 - x);
          double h0 = (d0 < 1.0e-10 || d1 < 1.0e-10) ? 1.0e-6 : 0.01 * sqrt(d0 / d1);
          h0 = fmin(h0, hmax);
#pragma unroll
          for (int r = 0; r < R; ++r) yt[r] = y[r] + f0[r] * h0;
          site(1, x + h0);
          if (status == 0) {
            LB_EFF_RHS(yt, f1, 1);
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int row = tid + W * r;
              const double sc = C.abstol + fabs(y[r]) * C.reltol;
              const double a0 = (f1[r] - f0[r]) / sc;
              eff_block_store(P, tid, W, r, L.red, 2, row < P.D ? a0 * a0 : 0.0);
            }
            __syncthreads();
            const double d2 = sqrt(eff_block_total(P, L.red, 2)) / h0;
            const double h1 = fmax(sqrt(d1), fabs(d2)) <= 1.0e-15 ? fmax(1.0e-6, fabs(h0) * 1.0e-3)
                                                                  : pow(0.01 / fmax(sqrt(d1), d2), 1.0 / 5.0);
            h = fmin(100.0 * fabs(h0), fmin(h1, hmax));
          }
        }
      }
      // k1 = f(t0, y0)
      if (status == 0) site(0, t0);
      if (status == 0) {
        LB_EFF_RHS(y, k[1], 0);
        rhs_calls = 1;
        if (!step_push) {
          while (si < C.n_saveat && fabs(C.saveat[si] - t0) <= time_tol(t0)) {
            push(C.saveat[si], y);
            ++si;
          }
        } else if (C.save_start) {
          push(t0, y);
        }
      }
      const double beta = 0.04, alpha = 0.2 - beta * 0.75, facc1 = 1.0 / 0.2, facc2 = 1.0 / 10.0;
      const double safety = 0.9, hmax = fabs(t1 - t0);
      double fac_old = 1.0e-4;
      bool reject = false, last = false;
      int64_t ns = 0, nst = 0, ia = 0;
      while (!last && status == 0) {
        if (ns > C.maxiters) { status = LB_TRAJ_MAXITERS; fail_x = x; break; }
        if (0.1 * fabs(h) <= DBL_EPSILON * fabs(x)) { status = LB_TRAJ_UNDERFLOW; fail_x = x; break; }
        if (x + 1.01 * h - t1 > 0.0) { h = t1 - x; last = true; }
        ++ns;
        // parameters at the six stage times (lanes 0-5), then the interval
        // lookups in call order (lane 0)
        if (tid < 6) {
          const int s = tid + 2;
          L.sp[4 * s + 2] = eff_params(P, C, traj, L.slots + s * P.n_slots, x + h * T::c(s), L.vm + tid * P.vm_depth,
                                       L.sp + 4 * s);
        }
        __syncthreads();
        if (tid == 0)
          for (int s = 2; s <= 7; ++s) eff_lookup(P, L, s);
        __syncthreads();
        if (L.misc[1] != 0) {
          status = L.misc[1];
          fail_x = L.ev[0];
          break;
        }
        const bool need_stiff = ((accepted + 1) % 1000 == 0) || ia > 0;
        double yt[R], yn[R];
#pragma unroll 1
        for (int s = 2; s <= 7; ++s) {
          switch (s) {
            case 2: eff_stage<R, SOLVER, 2>(h, y, k, yt); break;
            case 3: eff_stage<R, SOLVER, 3>(h, y, k, yt); break;
            case 4: eff_stage<R, SOLVER, 4>(h, y, k, yt); break;
            case 5: eff_stage<R, SOLVER, 5>(h, y, k, yt); break;
            case 6: eff_stage<R, SOLVER, 6>(h, y, k, yt); break;
            default:
              eff_stage<R, SOLVER, 7>(h, y, k, yn);
              if (need_stiff) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                  const double a0 = yn[r] - yt[r];
                  eff_block_store(P, tid, W, r, L.red, 3, a0 * a0);
                }
              }
#pragma unroll
              for (int r = 0; r < R; ++r) yt[r] = yn[r];
              break;
          }
          double ko[R];
          LB_EFF_RHS(yt, ko, s);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            switch (s) {
              case 2: k[2][r] = ko[r]; break;
              case 3: k[3][r] = ko[r]; break;
              case 4: k[4][r] = ko[r]; break;
              case 5: k[5][r] = ko[r]; break;
              case 6: k[6][r] = ko[r]; break;
              default: k[7][r] = ko[r]; break;
            }
          }
        }
        rhs_calls += 6;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = tid + W * r;
          double er = T::e(1) * k[1][r];
#pragma unroll
          for (int j = 2; j <= 7; ++j)
            if (T::e(j) != 0.0) er = fma(T::e(j), k[j][r], er);
          er *= h;
          const double sc = C.abstol + fmax(fabs(y[r]), fabs(yn[r])) * C.reltol;
          const double q = er / sc;
          eff_block_store(P, tid, W, r, L.red, 0, row < P.D ? q * q : 0.0);
          if (need_stiff) {
            const double a0 = k[7][r] - k[6][r];
            eff_block_store(P, tid, W, r, L.red, 2, a0 * a0);
          }
        }
        __syncthreads();
        const double err = sqrt(eff_block_total(P, L.red, 0) / (double)P.D);
        const int wl = tid & 63;
        const double pw = pow(wl < 32 ? err : fac_old, wl < 32 ? alpha : -beta);
        const double fac11 = __shfl(pw, 0, 64), fo = __shfl(pw, 32, 64);
        double fac = fac11 * fo;
        fac = fmax(facc2, fmin(facc1, fac / safety));
        double hn = h / fac;
        if (err <= 1.0) {
          fac_old = fmax(err, 1.0e-4);
          if (fabs(hn) > hmax) hn = hmax;
          if (reject) hn = fmin(fabs(hn), fabs(h));
          reject = false;
          ++accepted;
          if (accepted % 1000 == 0 || ia > 0) {
            const double n2 = eff_block_total(P, L.red, 2), dd2 = eff_block_total(P, L.red, 3);
            const double hl = dd2 > 0.0 ? h * sqrt(n2 / dd2) : 0.0;
            if (hl > 3.25) {
              ++ia;
              nst = 0;
              if (ia == 15) { status = LB_TRAJ_STIFF; fail_x = x; break; }
            } else {
              ++nst;
              if (nst == 6) ia = 0;
            }
          }
          const double xo = x;
          x += h;
          if (step_push) {
            push(x, yn);
          } else {
            while (si < C.n_saveat && C.saveat[si] <= x + time_tol(x)) {
              const double ts = C.saveat[si];
              if (ts >= xo - time_tol(xo)) {
                if (fabs(ts - xo) <= time_tol(xo)) {
                  push(ts, y);
                } else if (fabs(ts - x) <= time_tol(x)) {
                  push(ts, yn);
                } else {
                  double ds[R];
                  eff_dense<R, SOLVER>((ts - xo) / h, h, y, yn, k, ds);
                  push(ts, ds);
                }
              }
              ++si;
            }
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            y[r] = yn[r];
            k[1][r] = k[7][r];
          }
          if (last) break;
        } else {
          hn = h / fmin(facc1, fac11 / safety);
          reject = true;
          last = false;
          if (accepted >= 1) ++rejected;
        }
        h = hn;
      }
    }
#undef LB_EFF_RHS
    if (status == 0 && !step_push && si != C.n_saveat) {
      status = LB_TRAJ_SAVEAT;
      fail_x = (double)si;
    }
    if (status == 0 && !want_int && n_pushed > C.capacity) status = LB_TRAJ_CAPACITY;
    if (tid == 0) {
      if (want_int) {
        C.out_values[traj * C.capacity] = status == 0 ? integral : __builtin_nan("");
        if (C.out_times) C.out_times[traj * C.capacity] = last_t;
        C.out_counts[traj] = n_pushed > 0 ? 1 : 0;
      } else {
        C.out_counts[traj] = (int32_t)n_pushed;
      }
      C.out_stats[4 * traj + 0] = accepted;
      C.out_stats[4 * traj + 1] = rejected;
      C.out_stats[4 * traj + 2] = rhs_calls;
      C.out_stats[4 * traj + 3] = status;
      C.out_fail_t[traj] = fail_x;
    }
  }
}

// One effective RHS per workgroup (debug_effective_lindblad_rhs_py): the
// interval hint starts at 0 for every evaluation.
template <int R, int MAXW>
__global__ __launch_bounds__(MAXW, 1) void lb_eff_rhs_kernel(const DevEffPlan P, const DevEffCall C,
                                                             const DevEffRhsCall Q) {
  const int tid = threadIdx.x, W = blockDim.x;
  const LdsEff L = carve_eff(P);
  for (int64_t b = blockIdx.x; b < Q.batch; b += gridDim.x) {
    __syncthreads();
    if (tid == 0) {
      L.misc[1] = 0;
      L.misc[2] = 0;
      L.sp[2] = eff_params(P, C, 0, L.slots, Q.t[b], L.vm, L.sp);
      eff_lookup(P, L, 0);
    }
    __syncthreads();
    double y[R], k[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = tid + W * r;
      y[r] = row < P.D ? Q.y[b * (int64_t)P.D + row] : 0.0;
    }
    eff_rhs<R>(P, L.ybuf, y, k, L.sp, tid, W);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = tid + W * r;
      if (row < P.D) Q.out[b * (int64_t)P.D + row] = k[r];
    }
    if (tid == 0) Q.status[b] = L.misc[1];
  }
}

extern "C" void* lb_dev_eff_kernel(int id, int solver) {
#define E(i, r, w, v) \
  if (id == i) return solver == 1 ? (void*)&lb_eff_kernel<r, 1, w, v> : (void*)&lb_eff_kernel<r, 0, w, v>;
  LB_EFF_CONFIGS(E)
#undef E
  return nullptr;
}

extern "C" void* lb_dev_eff_rhs_kernel(int id) {
#define E(i, r, w, v) \
  if (id == i) return (void*)&lb_eff_rhs_kernel<r, w>;
  LB_EFF_CONFIGS(E)
#undef E
  return nullptr;
}
#endif  // LB_PART == 1

#if LB_PART == 0
__global__ void lb_queue_reset_kernel(int32_t* q) { *q = 0; }
#endif

// ---------------------------------------------------------------------------
// Kernel variants: (id, slots per thread R, largest workgroup, waves-per-SIMD
// hint).  The hint caps registers per lane (512 / hint); the host picks a
// variant with R * W >= plan slots (lb_host.cpp) and the autotuner measures.
// ---------------------------------------------------------------------------
// Variant table (id, slots per thread R, largest workgroup, waves-per-SIMD hint)
// in lb_variants.h.  The library compiles this file LB_NPART times; part
// LB_PART instantiates the variants with id % LB_NPART == LB_PART, so the
// parts build in parallel.
#ifndef LB_PART
#define LB_PART 0
#define LB_NPART 1
#endif
#define LB_CAT2(a, b) a##b
#define LB_CAT(a, b) LB_CAT2(a, b)
#define LB_PART_FN(name) LB_CAT(name##_p, LB_PART)

namespace {
template <int I, int RR, int WW, int VV>
void* solve_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 1 ? (void*)&lb_solve_kernel<RR, 1, WW, VV, 0> : (void*)&lb_solve_kernel<RR, 0, WW, VV, 0>;
  else
    return nullptr;
}
template <int I, int RR, int WW, int VV>
void* adjoint_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 1 ? (void*)&lb_adjoint_kernel<RR, 1, WW, VV> : (void*)&lb_adjoint_kernel<RR, 0, WW, VV>;
  else
    return nullptr;
}
template <int I, int RR, int WW, int VV>
void* fixed_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 2 ? (void*)&lb_fixed_kernel<RR, 0, WW, VV>
         : solver == 3 ? (void*)&lb_fixed_kernel<RR, 1, WW, VV>
                       : (void*)&lb_fixed_kernel<RR, 2, WW, VV>;
  else
    return nullptr;
}
// Stop-event forms: adaptive (solvers 0, 1) and fixed-step (2, 3, 4).
template <int I, int RR, int WW, int VV>
void* event_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 0   ? (void*)&lb_solve_kernel<RR, 0, WW, VV, 0, true>
         : solver == 1 ? (void*)&lb_solve_kernel<RR, 1, WW, VV, 0, true>
         : solver == 2 ? (void*)&lb_fixed_kernel<RR, 0, WW, VV, true>
         : solver == 3 ? (void*)&lb_fixed_kernel<RR, 1, WW, VV, true>
                       : (void*)&lb_fixed_kernel<RR, 2, WW, VV, true>;
  else
    return nullptr;
}
template <int I, int GG, int VV>
void* pair_solve_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 1 ? (void*)&lb_solve_pair_kernel<1, GG, VV> : (void*)&lb_solve_pair_kernel<0, GG, VV>;
  else
    return nullptr;
}
template <int I, int GG, int VV>
void* adjoint_wave_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 1 ? (void*)&lb_adjoint_wave_kernel<1, GG, VV> : (void*)&lb_adjoint_wave_kernel<0, GG, VV>;
  else
    return nullptr;
}
template <int I, int GG, int VV>
void* wave_solve_ptr(int solver) {
  if constexpr (I % LB_NPART == LB_PART)
    return solver == 1 ? (void*)&lb_solve_kernel<1, 1, 64, VV, GG> : (void*)&lb_solve_kernel<1, 0, 64, VV, GG>;
  else
    return nullptr;
}
template <int I, int VV>
void* wave_rhs_ptr() {
  if constexpr (I % LB_NPART == LB_PART)
    return (void*)&lb_rhs_kernel<1, 64, VV>;
  else
    return nullptr;
}
template <int I, int RR, int WW, int VV>
void* rhs_ptr() {
  if constexpr (I % LB_NPART == LB_PART)
    return (void*)&lb_rhs_kernel<RR, WW, VV>;
  else
    return nullptr;
}
}  // namespace

extern "C" void* LB_PART_FN(lb_dev_solve_kernel)(int id, int solver) {
#define X(i, r, w, v) \
  if (id == i) return solve_ptr<i, r, w, v>(solver);
  LB_CONFIGS(X)
#undef X
#define Y(i, g, v) \
  if (id == i) return wave_solve_ptr<i, g, v>(solver);
  LB_WAVE_CONFIGS(Y)
#undef Y
#define Q(i, g, v) \
  if (id == i) return pair_solve_ptr<i, g, v>(solver);
  LB_PAIR_CONFIGS(Q)
#undef Q
  return nullptr;
}

extern "C" void* LB_PART_FN(lb_dev_fixed_kernel)(int id, int solver) {
#define F(i, r, w, v) \
  if (id == i) return fixed_ptr<i, r, w, v>(solver);
  LB_FIXED_CONFIGS(F)
#undef F
  return nullptr;
}

extern "C" void* LB_PART_FN(lb_dev_event_kernel)(int id, int solver) {
#define E(i, r, w, v) \
  if (id == i) return event_ptr<i, r, w, v>(solver);
  LB_EVENT_CONFIGS(E)
#undef E
  return nullptr;
}

extern "C" void* LB_PART_FN(lb_dev_adjoint_kernel)(int id, int solver) {
#define Z(i, r, w, v) \
  if (id == i) return adjoint_ptr<i, r, w, v>(solver);
  LB_ADJ_CONFIGS(Z)
#undef Z
#define V(i, g, v) \
  if (id == i) return adjoint_wave_ptr<i, g, v>(solver);
  LB_ADJ_WAVE_CONFIGS(V)
#undef V
  return nullptr;
}

extern "C" void* LB_PART_FN(lb_dev_rhs_kernel)(int id) {
#define X(i, r, w, v) \
  if (id == i) return rhs_ptr<i, r, w, v>();
  LB_CONFIGS(X)
#undef X
#define Y(i, g, v) \
  if (id == i) return wave_rhs_ptr<i, v>();
  LB_WAVE_CONFIGS(Y)
#undef Y
  return nullptr;
}

#if LB_PART == 0
extern "C" void* lb_dev_queue_reset_kernel(void) { return (void*)&lb_queue_reset_kernel; }
#endif
