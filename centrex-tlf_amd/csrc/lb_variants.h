// Compiled solver variants: X(id, slots per thread R, largest workgroup,
// waves-per-SIMD hint).  The hint caps the registers per lane at 512 / hint;
// the host picks a variant whose R * W covers the plan's slots.
#pragma once

// Two shapes gave wrong results that no other variant showed, deterministic
// and with no cause found in the kernel: (4, 256, 1), once entry 6, on the
// full n=39 plan, and (12, 256, 1) and (12, 256, 2), once entry 10, on the
// restricted n=72 plan (12 slots per lane, one wave, about 1,300 spilled
// VGPRs).  Neither is compiled; every entry is checked bitwise against the
// default launch on the n=72 plan.
#define LB_CONFIGS(X)                                                                            \
  X(0, 1, 1024, 4) X(1, 2, 512, 4) X(2, 2, 512, 2) X(3, 3, 1024, 4) X(4, 4, 256, 4)                \
  X(5, 4, 256, 2) X(6, 3, 512, 2) X(7, 6, 512, 2) X(8, 6, 512, 1) X(9, 8, 128, 2) X(10, 6, 256, 2)

// Single-wave variants for plans of at most 64 elements: Y(id, gather rows
// held in registers, waves-per-SIMD hint); one slot per lane, 64 threads.
#define LB_WAVE_CONFIGS(Y) Y(11, 8, 4) Y(12, 8, 2) Y(13, 16, 2) Y(14, 12, 2)

// Paired variants for plans of at most 32 elements: two trajectories per wave
// (one per 32-lane half); Q(id, gather rows in registers, waves-per-SIMD
// hint).  Only for persistent calls that push at every accepted step with a
// given initial step (no trace, save plan or replay; a tape may be recorded).
#define LB_PAIR_CONFIGS(Q) Q(15, 8, 2) Q(16, 12, 2)

// Fixed-step kernel variants (fixed_dopri5 / fixed_rk2 / fixed_rk4):
// F(id, slots per thread R, largest workgroup, waves-per-SIMD hint).
#define LB_FIXED_CONFIGS(F) F(0, 1, 1024, 2) F(1, 2, 512, 2) F(2, 4, 256, 2) F(3, 8, 256, 2)
// Stop-event forms (adaptive and fixed-step, compiled apart so the event code
// leaves the other kernels' register allocation alone): one slot per lane up
// to 1024 slots, six per lane up to 3072.
#define LB_EVENT_CONFIGS(E) E(0, 1, 1024, 4) E(1, 6, 512, 1)
// Effective-Hamiltonian integrator: rows of the real state per lane, largest
// workgroup, waves-per-SIMD hint.
#define LB_EFF_CONFIGS(E) E(0, 1, 1024, 4) E(1, 4, 512, 2) E(2, 8, 512, 1)

// Adjoint kernel variants: Z(id, slots per thread R, largest workgroup, waves-per-SIMD hint).
#define LB_ADJ_CONFIGS(Z) Z(0, 1, 1024, 2) Z(1, 2, 512, 2) Z(2, 4, 256, 2) Z(3, 8, 256, 2)

// Single-block adjoint variants for plans of one 64-slot block (lane = slot,
// several waves per trajectory): V(id, gather rows held in registers by
// the recomputation, waves-per-SIMD hint), in increasing rows (the host takes
// the first that covers the plan).
#define LB_ADJ_WAVE_CONFIGS(V) V(4, 12, 2) V(5, 16, 2)

#define LB_NPARTS_BUILT 4
