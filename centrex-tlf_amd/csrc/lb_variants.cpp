// Host view of the variant table and of the kernels compiled in the parts of
// lb_kernels.hip (one part per object file, LB_NPARTS_BUILT in total).
#include "lb_variants.h"

extern "C" void* lb_dev_solve_kernel_p0(int, int);
extern "C" void* lb_dev_solve_kernel_p1(int, int);
extern "C" void* lb_dev_solve_kernel_p2(int, int);
extern "C" void* lb_dev_solve_kernel_p3(int, int);
extern "C" void* lb_dev_fixed_kernel_p0(int, int);
extern "C" void* lb_dev_fixed_kernel_p1(int, int);
extern "C" void* lb_dev_fixed_kernel_p2(int, int);
extern "C" void* lb_dev_fixed_kernel_p3(int, int);
extern "C" void* lb_dev_event_kernel_p0(int, int);
extern "C" void* lb_dev_event_kernel_p1(int, int);
extern "C" void* lb_dev_event_kernel_p2(int, int);
extern "C" void* lb_dev_event_kernel_p3(int, int);
extern "C" void* lb_dev_adjoint_kernel_p0(int, int);
extern "C" void* lb_dev_adjoint_kernel_p1(int, int);
extern "C" void* lb_dev_adjoint_kernel_p2(int, int);
extern "C" void* lb_dev_adjoint_kernel_p3(int, int);
extern "C" void* lb_dev_rhs_kernel_p0(int);
extern "C" void* lb_dev_rhs_kernel_p1(int);
extern "C" void* lb_dev_rhs_kernel_p2(int);
extern "C" void* lb_dev_rhs_kernel_p3(int);

static_assert(LB_NPARTS_BUILT == 4, "lb_variants.cpp lists four kernel parts");

extern "C" int lb_dev_variant_count(void) {
  int n = 0;
#define X(i, r, w, v) ++n;
  LB_CONFIGS(X)
#undef X
#define Y(i, g, v) ++n;
  LB_WAVE_CONFIGS(Y)
#undef Y
#define Q(i, g, v) ++n;
  LB_PAIR_CONFIGS(Q)
#undef Q
  return n;
}

extern "C" int lb_dev_variant(int id, int* R, int* maxw, int* waves) {
#define X(i, r, w, v)                                                                           \
  if (id == i) { *R = r; *maxw = w; *waves = v; return 1; }
  LB_CONFIGS(X)
#undef X
#define Y(i, g, v)                                                                              \
  if (id == i) { *R = 1; *maxw = 64; *waves = v; return 1; }
  LB_WAVE_CONFIGS(Y)
#undef Y
#define Q(i, g, v)                                                                              \
  if (id == i) { *R = 1; *maxw = 64; *waves = v; return 1; }
  LB_PAIR_CONFIGS(Q)
#undef Q
  return 0;
}

// 1 for the paired (two trajectories per wave) variants.
extern "C" int lb_dev_variant_pair(int id) {
#define Q(i, g, v) \
  if (id == i) return 1;
  LB_PAIR_CONFIGS(Q)
#undef Q
  return 0;
}

// Gather rows a single-wave variant holds in registers (0: multi-wave variant).
extern "C" int lb_dev_variant_rows(int id) {
#define Y(i, g, v) \
  if (id == i) return g;
  LB_WAVE_CONFIGS(Y)
#undef Y
#define Q(i, g, v) \
  if (id == i) return g;
  LB_PAIR_CONFIGS(Q)
#undef Q
  return 0;
}

extern "C" void* lb_dev_solve_kernel(int id, int solver) {
  void* (*parts[])(int, int) = {lb_dev_solve_kernel_p0, lb_dev_solve_kernel_p1, lb_dev_solve_kernel_p2,
                                lb_dev_solve_kernel_p3};
  return id < 0 ? nullptr : parts[id % LB_NPARTS_BUILT](id, solver);
}

extern "C" void* lb_dev_rhs_kernel(int id) {
  void* (*parts[])(int) = {lb_dev_rhs_kernel_p0, lb_dev_rhs_kernel_p1, lb_dev_rhs_kernel_p2,
                           lb_dev_rhs_kernel_p3};
  return id < 0 ? nullptr : parts[id % LB_NPARTS_BUILT](id);
}

// Adjoint variants: Z(id, R, largest workgroup, waves hint), then the
// single-wave V(id, gather rows, waves hint) with *rows = gather rows (0 for
// the multi-wave form).
extern "C" int lb_dev_adjoint_variant(int id, int* R, int* maxw, int* rows) {
#define Z(i, r, w, v)                                                                           \
  if (id == i) { *R = r; *maxw = w; *rows = 0; return 1; }
  LB_ADJ_CONFIGS(Z)
#undef Z
#define V(i, g, v)                                                                              \
  if (id == i) { *R = 1; *maxw = 64; *rows = g; return 1; }
  LB_ADJ_WAVE_CONFIGS(V)
#undef V
  return 0;
}

extern "C" void* lb_dev_adjoint_kernel(int id, int solver) {
  void* (*parts[])(int, int) = {lb_dev_adjoint_kernel_p0, lb_dev_adjoint_kernel_p1, lb_dev_adjoint_kernel_p2,
                                lb_dev_adjoint_kernel_p3};
  return id < 0 ? nullptr : parts[id % LB_NPARTS_BUILT](id, solver);
}

// Fixed-step variants: F(id, R, largest workgroup, waves hint).
extern "C" int lb_dev_fixed_variant(int id, int* R, int* maxw) {
#define F(i, r, w, v)                                                                           \
  if (id == i) { *R = r; *maxw = w; return 1; }
  LB_FIXED_CONFIGS(F)
#undef F
  return 0;
}

extern "C" int lb_dev_eff_variant(int id, int* R, int* maxw) {
#define E(i, r, w, v)                                                                           \
  if (id == i) { *R = r; *maxw = w; return 1; }
  LB_EFF_CONFIGS(E)
#undef E
  return 0;
}

extern "C" int lb_dev_event_variant(int id, int* R, int* maxw) {
#define E(i, r, w, v)                                                                           \
  if (id == i) { *R = r; *maxw = w; return 1; }
  LB_EVENT_CONFIGS(E)
#undef E
  return 0;
}

extern "C" void* lb_dev_event_kernel(int id, int solver) {
  void* (*parts[])(int, int) = {lb_dev_event_kernel_p0, lb_dev_event_kernel_p1, lb_dev_event_kernel_p2,
                                lb_dev_event_kernel_p3};
  return id < 0 ? nullptr : parts[id % LB_NPARTS_BUILT](id, solver);
}

extern "C" void* lb_dev_fixed_kernel(int id, int solver) {
  void* (*parts[])(int, int) = {lb_dev_fixed_kernel_p0, lb_dev_fixed_kernel_p1, lb_dev_fixed_kernel_p2,
                                lb_dev_fixed_kernel_p3};
  return id < 0 ? nullptr : parts[id % LB_NPARTS_BUILT](id, solver);
}
