#!/usr/bin/env python3
"""Benchmark: batched persistent GPU solve of the n=39 TlF 64x64 grid scan.

Workload (one "step"): X(J=0,1,2) + B(J=1,F1'=1/2,F'=1), n = 39, one z laser
with a Gaussian Rabi envelope Omega(t) = Omega0 exp(-(t-t0)^2/(2 sigma^2)),
sigma = 3/Gamma, t0 = 10/Gamma, t in [0, 20/Gamma], Gamma = 2 pi 1.56 MHz;
grid delta0 = linspace(-5, 5, 64) Gamma (slowest axis) x Omega0 =
geomspace(0.1, 10, 64) Gamma -> 4096 trajectories, Dopri5, rtol 1e-8,
atol 1e-10, dt 1e-8.  Each trajectory returns its final populations and the
photon count (trapezoid over accepted steps) from one fused launch.  The
initial state, the plan tables and the scan axes are resident on the GPU
before the timed region.

The solver integrates only the density-matrix elements reachable from the
initial state through the RHS program (23 of the 780 upper-triangle elements
here; all others stay exactly zero), two trajectories per wave (one per
32-lane half).

Also reported: the per-trajectory ("plain") GPU path on a seeded sample, the
CPU oracle (this project's C restatement of the reference's Rust/rayon path,
which integrates every element) timed on the host cores on a bounded seeded
sample, a parity check of the sample against the oracle, and the FP64
roofline of the solver kernel: `achieved` counts the reference algorithm's
FLOPs (rhs_calls * F_rhs + attempted * dim * F_rk over the full plan), and
`executed_*` the FLOPs of the restricted tables the kernel actually runs.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --copy-calibration   # plain 1 GiB copy kernel (HBM counter scale)
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT / "centrex-tlf_amd"), str(ROOT)):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
F_RK = {"dopri5": 65, "tsit5": 67}  # algorithmic FLOPs per packed element per attempted step
TRAFFIC_FILE = ROOT / "profiles" / "traffic_lb_solve_pair_kernel_n39_grid4096.json"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# LDS of a CU: 128 bytes per clock (the co-bound's peak: 256 CUs at 2.4 GHz)
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12


def lds_bytes_per_rhs(prepared) -> int:
    """Algorithmic LDS bytes of one RHS of the full plan: 8 bytes per term whose
    input is a real (diagonal) element, 16 per complex one, plus 8*dim for
    publishing the stage vector that the gathers read."""
    import numpy as np

    from centrex_tlf_amd.lindblad.state_layout import upper_offsets

    plan = prepared.expanded_rhs_plan
    n = int(plan["n"])
    inputs = np.asarray(plan["input_indices"], dtype=np.int64)
    diagonal = np.isin(inputs, upper_offsets(n))
    return int(np.where(diagonal, 8, 16).sum()) + 8 * n * n


def algorithmic_flops(stats, flops_per_rhs: int, dim: int, solver: str) -> float:
    """Sum over trajectories of rhs_calls*F_rhs + attempted_steps*dim*F_rk.

    attempted steps = (rhs_calls - 1) / 6 (one FSAL call, then six per attempt).
    """
    import numpy as np

    rhs = stats[:, 2].astype(np.float64)
    attempted = (rhs - 1.0) / 6.0
    return float(np.sum(rhs * flops_per_rhs + attempted * dim * F_RK[solver]))


def uneven_leg(steps: int, warmup: int, cpu_seconds: float) -> dict:
    """Uneven batch in one call: 512 n=39 + 512 n=72 trajectories.

    n=72: X(J=0..3) + B(J=1, F1'=3/2, F'=1,2), two lasers whose polarisation
    switches x <-> y with a 1 MHz square wave, microwave J=1-J=2 at 0.5 Gamma.
    Every trajectory has its own width sigma ~ N(5, 1.5)/Gamma (clipped above
    1/Gamma), centre sigma and end time 2 sigma, and a detuning ~ U[-2, 2]
    Gamma; rtol 1e-8, atol 1e-10.  The launch parameters come from the
    autotuner (timed on this batch); the untuned launch (each group filling
    the device in turn) is reported beside it.  Dopri5 is the headline, Tsit5
    is timed too.  The CPU oracle runs a seeded sample of both systems on the
    host cores; its populations are compared with the GPU's.
    """
    import numpy as np
    import torch

    from centrex_tlf_amd.lindblad import UnevenGroup, autotune, prepare_lindblad_problem, solve_uneven_batch
    from oracle.lb_oracle import Oracle, host_threads
    from tests.systems import GAMMA, tlf_n39, tlf_n72

    tol = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
    names = ["sigma", "t_center", "delta0"]
    rng = np.random.default_rng(20261202 + 2)
    groups = []
    for build in (tlf_n39, tlf_n72):
        system, params, rho0, _, weights = build()
        prepared = prepare_lindblad_problem(system, params, backend="gpu")
        sigma = np.clip(rng.normal(5.0, 1.5, 512), 1.0, None) / GAMMA
        delta = rng.uniform(-2.0, 2.0, 512) * GAMMA
        pb = np.stack([sigma, sigma, delta], axis=1).astype(complex)
        groups.append(UnevenGroup(prepared, rho0, 2.0 * sigma, names, pb, weights))
    n_traj = sum(len(g.t_end) for g in groups)
    out: dict = {"metric": "uneven batch trajectories/s", "unit": "trajectories/s", "trajectories": n_traj,
                 "config": "512 tlf_n39 + 512 tlf_n72, per-trajectory sigma and t_end = 2 sigma, rtol 1e-8"}

    def timed(fn):
        for _ in range(max(1, warmup)):
            res = fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            res = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, res

    for solver in ("dopri5", "tsit5"):
        t_tune = time.perf_counter()
        tun = autotune(groups, solver=solver, repeats=1, **tol)
        t_tune = time.perf_counter() - t_tune
        tuned_s, res = timed(lambda: solve_uneven_batch(groups, solver=solver, tuning=tun, **tol))
        plain_s, _ = timed(lambda: solve_uneven_batch(groups, solver=solver, tuning=None, longest_first=False,
                                                      **tol))
        steps_all = [r.stats[:, 0] for r in res]
        log(f"bench: uneven {solver}: tuned {tuned_s * 1e3:.1f} ms ({tun.mode}, variants {tun.variants}, "
            f"workgroups {tun.workgroups}), untuned {plain_s * 1e3:.1f} ms, tuning took {t_tune:.1f} s")
        out[solver] = {
            "value": n_traj / tuned_s,
            "ms_per_batch": tuned_s * 1e3,
            "untuned_ms_per_batch": plain_s * 1e3,
            "tuning": {"mode": tun.mode, "variants": tun.variants, "threads": tun.threads,
                       "workgroups": tun.workgroups, "seconds": t_tune},
            "accepted_steps_min_max": {"n39": [int(steps_all[0].min()), int(steps_all[0].max())],
                                       "n72": [int(steps_all[1].min()), int(steps_all[1].max())]},
        }
        if solver == "dopri5":
            out["value"] = n_traj / tuned_s
            pops = [r.populations for r in res]
            # roofline of the batch: the reference algorithm's FLOPs (full
            # plans) and the FLOPs of the restricted tables the kernels run,
            # over the grouped call's time; the n=72 group sets that time
            roof = {"bound": "fp64_valu", "latency_bound": True, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "seconds": tuned_s, "groups": {}}
            alg_all = exe_all = 0.0
            for grp, r, label in zip(groups, res, ("n39", "n72")):
                plan = grp.prepared.device_plan()
                y0 = grp.prepared.layout.pack(grp.rho0)
                info, rinfo = plan.info, plan.support_info(plan.support_of(y0))
                st = r.stats.cpu().numpy() if hasattr(r.stats, "cpu") else np.asarray(r.stats)
                alg = algorithmic_flops(st, info["flops_per_rhs"], info["dim"], solver)
                exe = algorithmic_flops(st, rinfo["flops_per_rhs"], rinfo["dim"], solver)
                roof["groups"][label] = {"flops": alg, "executed_flops": exe, "elements": int(rinfo["upper_len"]),
                                         "f_rhs_full": int(info["flops_per_rhs"]),
                                         "f_rhs_restricted": int(rinfo["flops_per_rhs"])}
                alg_all += alg
                exe_all += exe
            roof["achieved"] = alg_all / tuned_s / 1e12
            roof["frac"] = roof["achieved"] / FP64_PEAK_TFLOPS
            roof["executed_frac"] = exe_all / tuned_s / 1e12 / FP64_PEAK_TFLOPS
            out["roofline"] = roof
    # ---- CPU oracle on all host cores: a seeded sample of each system solved
    # as one batch (per-trajectory end times), Dopri5; the mixed batch's rate
    # follows from the two measured rates
    cores = host_threads()
    cpu = {}
    errs = []
    for k, (grp, label, per_core) in enumerate(zip(groups, ("n39", "n72"), (8, 1))):
        oracle = Oracle(grp.prepared)
        y0 = grp.prepared.layout.pack(grp.rho0)
        take = np.sort(rng.choice(len(grp.t_end), size=min(len(grp.t_end), per_core * cores), replace=False))
        t_c = time.perf_counter()
        r = oracle.solve_batch(y0, (0.0, float(grp.t_end[take].max())), output="populations",
                               output_when="final", parameter_slots=names, parameter_batch=grp.parameter_batch[take],
                               t1_per_traj=grp.t_end[take], threads=cores, **tol)
        cpu[label] = len(take) / (time.perf_counter() - t_c)  # trajectories/s on all cores
        errs.append(float(np.abs(r["final"] - pops[k][take]).max()))
        cpu[label + "_sample"] = int(len(take))
    cpu_rate = n_traj / (512 / cpu["n39"] + 512 / cpu["n72"])
    out["cpu_baseline"] = {"value": cpu_rate, "unit": "trajectories/s", "cores": cores, "kind": "port",
                           "sample": f"oracle on {cores} threads, measured: {cpu['n39_sample']} seeded n39 "
                                     f"trajectories at {cpu['n39']:.1f}/s and {cpu['n72_sample']} n72 at "
                                     f"{cpu['n72']:.2f}/s, combined for the 512 + 512 batch"}
    out["speedup_vs_cpu"] = out["value"] / cpu_rate
    out["parity"] = {"trajectories": cpu["n39_sample"] + cpu["n72_sample"], "max_abs_population_error": max(errs),
                     "bar": "1e-7 absolute"}
    return out


def fit_leg(per_group: int, iterations: int, t_scale: float, graph: bool = True) -> dict:
    """Batched gradient-based fit: per_group n=39 (x/y) + per_group n=72 (theta)
    trajectories, sigma ~ N(5, 1.5)/Gamma clipped above 1/Gamma (times t_scale),
    t_center = sigma, t_end = 2 sigma; fit per-trajectory (delta0, Omega0,
    theta) by Adam (lr 0.05 in units of Gamma, Gamma, 1) to photon counts of
    seeded true parameters, starting from perturbed ones; Dopri5, rtol 1e-8,
    checkpoints every 16 accepted steps.  With ``graph`` every iteration after
    the first is one replay of a HIP graph holding both systems' forward
    solves, adjoints and the Adam step (statuses read once, after the last
    replay).  Reports gradients/s, the tape memory, the loss curve, and the
    final populations of a sample (4 trajectories per system, the shortest)
    against the CPU oracle at rtol 1e-8.
    """
    import numpy as np
    import torch

    from centrex_tlf_amd.lindblad import prepare_lindblad_problem
    from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount
    from centrex_tlf_amd.lindblad.fit import FitGroup, fit_photon_counts
    from oracle.lb_oracle import Oracle
    from tests.systems import GAMMA, tlf_n39_xy, tlf_n72_theta

    names = ["delta0", "Omega0", "theta"]
    tol = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
    rng = np.random.default_rng(20261202 + 5)
    groups = []
    for build in (tlf_n39_xy, tlf_n72_theta):
        system, params, rho0, _, weights = build()
        prepared = prepare_lindblad_problem(system, params, backend="gpu")
        sigma = np.clip(rng.normal(5.0, 1.5, per_group), 1.0, None) / GAMMA * t_scale
        t_end = 2.0 * sigma
        true = np.stack([rng.uniform(-3, 3, per_group) * GAMMA, rng.uniform(0.3, 3, per_group) * GAMMA,
                         rng.uniform(0, np.pi, per_group)], axis=1)
        fixed = {"sigma": sigma.astype(complex), "t_center": sigma.astype(complex)}
        eng = DifferentiablePhotonCount(prepared, names, weights, (0.0, t_end), fixed=fixed, **tol)
        targets, _ = eng.forward(true, prepared.layout.pack(rho0))
        initial = true * np.asarray([1.0, 1.2, 1.0]) + np.asarray([0.3 * GAMMA, 0.0, 0.2])
        groups.append(FitGroup(prepared, rho0, t_end, targets.cpu().numpy(), initial, weights, fixed))
    log(f"bench: fit of {2 * per_group} trajectories x {iterations} iterations")
    def progress(it):
        if it % 5 == 0:
            torch.cuda.synchronize()
            log(f"bench: fit iteration {it} done")

    res = fit_photon_counts(groups, names, scales=[GAMMA, GAMMA, 1.0], iterations=iterations, lr=0.05,
                            graph=graph, on_replay=progress, **tol)
    log(f"bench: fit done in {res.seconds:.1f} s")
    # final populations of a sample against the oracle
    errs = []
    for grp, p in zip(groups, res.parameters):
        plan = grp.prepared.device_plan()
        y0 = grp.prepared.layout.pack(grp.rho0)
        slots = [grp.prepared.slot_names.index(m) for m in names + ["sigma", "t_center"]]
        vals = np.concatenate([p, np.stack([grp.fixed["sigma"].real, grp.fixed["t_center"].real], axis=1)], axis=1)
        pops = plan.solve(y0, 0.0, grp.t_end, output_kind="populations", override_slots=slots,
                          override_values=vals.astype(complex), **tol)[0].cpu().numpy()
        oracle = Oracle(grp.prepared)
        for b in np.argsort(grp.t_end)[:4]:
            r = oracle.solve_batch(y0, (0.0, float(grp.t_end[b])), output="populations", output_when="final",
                                   parameter_slots=names + ["sigma", "t_center"],
                                   parameter_batch=vals[b:b + 1].astype(complex), threads=1, **tol)
            errs.append(float(np.abs(r["final"][0] - pops[b]).max()))
    return {
        "metric": "fit gradients/s",
        "value": res.gradients_per_second,
        "unit": "trajectory gradients (delta0, Omega0, theta) per second, forward + adjoint + Adam",
        "trajectories": 2 * per_group,
        "iterations": iterations,
        "seconds": res.seconds,
        # the replayed iterations run forward and adjoint in one graph: no split
        "forward_seconds": None if graph else res.forward_seconds,
        "adjoint_seconds": None if graph else res.adjoint_seconds,
        "loss_first_last": [res.loss[0], res.loss[-1]],
        "loss": res.loss,
        "graph_replayed": graph,
        "tape_bytes_peak": res.tape_bytes,
        "final_population_error_vs_oracle": max(errs),
        "config": f"{per_group} tlf_n39_xy + {per_group} tlf_n72_theta, sigma scale {t_scale}, rtol 1e-8",
    }


def adjoint_leg(steps: int, warmup: int) -> dict:
    """Gradients of the photon-count loss on 256 n=39 trajectories.

    System: the n=39 set with the laser polarisation split into x and y
    (PX = cos theta, PY = sin theta), Gaussian envelope of width sigma.
    Parameters per trajectory (seeded): delta0 ~ U[-3, 3] Gamma, Omega0 ~
    U[0.3, 3] Gamma, theta ~ U[0, pi], sigma ~ U[2, 6] / Gamma; targets ~
    U[0, 1]; the initial state a seeded coherent superposition of the
    X(J=0) levels.  Loss sum_b (N_b - target_b)^2, Dopri5 rtol 1e-8 atol 1e-10,
    checkpoints every 16 accepted steps.  One step = forward with tape +
    adjoint backward through ``torch.autograd``; the forward alone is the
    ordinary batched photon-count solve of the same trajectories.
    """
    import numpy as np
    import torch

    from centrex_tlf_amd.lindblad import prepare_lindblad_problem
    from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount, photon_count
    from tests.systems import GAMMA, tlf_n39_xy

    names = ["delta0", "Omega0", "theta", "sigma"]
    tol = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
    n = 256
    rng = np.random.default_rng(20261202 + 4)
    p_np = np.stack([rng.uniform(-3, 3, n) * GAMMA, rng.uniform(0.3, 3, n) * GAMMA, rng.uniform(0, np.pi, n),
                     rng.uniform(2, 6, n) / GAMMA], axis=1)
    dev = torch.device("cuda", 0)
    target = torch.as_tensor(rng.uniform(0, 1, n), device=dev)
    system, params, _, t_span, weights = tlf_n39_xy()
    # initial state: a seeded coherent superposition of the X(J=0) levels (an
    # isotropic mixture makes dN/dtheta vanish by symmetry)
    ground = system.indices(electronic="X", J=0)
    psi = np.zeros(system.n_states, dtype=np.complex128)
    crng = np.random.default_rng(7)
    psi[ground] = crng.normal(size=len(ground)) + 1j * crng.normal(size=len(ground))
    psi /= np.linalg.norm(psi)
    rho0 = np.outer(psi, psi.conj())
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    engine = DifferentiablePhotonCount(prepared, names, weights, t_span, ckpt_every=16, **tol)
    y0 = torch.as_tensor(prepared.layout.pack(rho0), device=dev)
    p = torch.as_tensor(p_np, device=dev)
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index(m) for m in names]
    pv = p.to(torch.complex128)

    def forward_only(events=None):
        return plan.solve(y0, *t_span, output_kind="integral", weights=weights, override_slots=slots,
                          override_values=pv, events=events, **tol)

    def grad_step():
        q = p.clone().requires_grad_(True)
        counts = photon_count(engine, q, y0)
        loss = ((counts - target) ** 2).sum()
        loss.backward()
        return counts, q.grad

    def timed(fn):
        for _ in range(max(1, warmup)):
            out = fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, out

    def kernel_ms(fn):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for e in ev:
            fn(e)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev]))

    fwd_s, (counts_ref, stats, _) = timed(forward_only)
    grad_s, (counts, grad) = timed(grad_step)
    # parts of one gradient step, timed apart
    tape_s, (_, tape) = timed(lambda: engine.forward(p, y0))
    bwd_s, _ = timed(lambda: engine.gradient(p, y0, tape))
    # the kernels alone (HIP events around each launch)
    fwd_kms = kernel_ms(lambda e: forward_only(events=e))
    tape_kms = kernel_ms(lambda e: engine.forward(p, y0, events=e))
    bwd_kms = kernel_ms(lambda e: engine.gradient(p, y0, tape, events=e))
    # the gradient against central differences (h = 1e-6 |p|) of the GPU
    # forward on all 256 trajectories: relative to the batch's largest
    # |dN/dp_m| per parameter; a trajectory whose attempt count changes under
    # +-h is a jump of the adaptive map, reported and left out
    g_all = engine.gradient(p, y0, tape).cpu().numpy()
    n0 = tape.n.cpu().numpy()
    fd = np.zeros_like(p_np)
    same = np.ones(n, dtype=bool)
    for m in range(4):
        h = 1e-6 * np.abs(p_np[:, m])
        hi, lo = p_np.copy(), p_np.copy()
        hi[:, m] += h
        lo[:, m] -= h
        (ch, th), (cl, tl) = engine.forward(hi, y0), engine.forward(lo, y0)
        fd[:, m] = (ch.cpu().numpy() - cl.cpu().numpy()) / (2 * h)
        same &= (th.n.cpu().numpy() == n0) & (tl.n.cpu().numpy() == n0)
    scale = np.abs(fd[same]).max(axis=0)
    err = (np.abs(g_all - fd)[same] / scale).max(axis=0)
    frozen = engine.gradient(p, y0, tape, mode="frozen").cpu().numpy()
    err_frozen = (np.abs(frozen - fd)[same] / scale).max(axis=0)
    # the frozen-step derivative against central differences of a replay of
    # the recorded step sizes, on 8 trajectories
    k = 8
    pk = p_np[:k]
    _, tape_k = engine.forward(pk, y0)
    gk = engine.gradient(pk, y0, tape_k, mode="frozen").cpu().numpy()
    fdr = np.zeros_like(pk)
    for m in range(4):
        h = 1e-6 * np.abs(pk[:, m])
        hi, lo = pk.copy(), pk.copy()
        hi[:, m] += h
        lo[:, m] -= h
        fdr[:, m] = (engine.replay(hi, y0, tape_k).cpu().numpy() - engine.replay(lo, y0, tape_k).cpu().numpy()) / (2 * h)
    err_replay = (np.abs(gk - fdr) / np.abs(fdr).max(axis=0)).max(axis=0)
    st = stats.cpu().numpy()
    log(f"bench: adjoint {n / grad_s:.1f} grads/s; forward {fwd_s * 1e3:.2f} ms, "
        f"tape forward {tape_s * 1e3:.2f} ms, backward {bwd_s * 1e3:.2f} ms; kernels {fwd_kms:.2f} / "
        f"{tape_kms:.2f} / {bwd_kms:.2f} ms")
    return {
        "metric": "adjoint grads/s",
        "value": n / grad_s,
        "unit": "full gradients (4 parameters) per second",
        "trajectories": n,
        "parameters": names,
        "ms_per_step": grad_s * 1e3,
        "forward_ms": fwd_s * 1e3,
        "tape_forward_ms": tape_s * 1e3,
        "backward_ms": bwd_s * 1e3,
        "backward_over_forward": bwd_s / fwd_s,
        "gradient_step_over_forward": grad_s / fwd_s,
        "kernel_ms": {"forward": fwd_kms, "forward_with_tape": tape_kms, "adjoint": bwd_kms},
        "kernel_adjoint_over_forward": bwd_kms / fwd_kms,
        "kernel_gradient_over_forward": (tape_kms + bwd_kms) / fwd_kms,
        "accepted_steps_min_max": [int(st[:, 0].min()), int(st[:, 0].max())],
        "checkpoint_every": 16,
        "tape_bytes": int(tape.nbytes),
        "counts_max_abs_diff_vs_forward": float((counts.detach() - counts_ref[:, 0]).abs().max().item()),
        "fd_error": [float(e) for e in err],
        "fd_error_definition": "per parameter, max over trajectories of |adjoint - central FD of the GPU forward| "
                               "/ max |FD|; h = 1e-6 |p|; bar 1e-5",
        "fd_trajectories": int(same.sum()),
        "fd_trajectories_with_step_count_change": int((~same).sum()),
        "frozen_step_vs_fd_error": [float(e) for e in err_frozen],
        "frozen_step_vs_replay_fd_error": [float(e) for e in err_replay],
        "config": "tlf_n39_xy, 256 seeded trajectories, coherent X(J=0) initial state, Dopri5 rtol 1e-8 atol 1e-10",
    }


def _effective_synthetic_plan(n: int, seed: int = 7, density: float = 0.02, n_grid: int = 6):
    """A synthetic effective-Hamiltonian plan: random sparse split-real
    operators (spectral radius ~1, the combined one damped) on a field grid,
    field x(t) = x0 + v t, Rabi rate a Gaussian pulse."""
    import numpy as np

    from centrex_tlf_amd.lindblad import LindbladParameters, Time, gaussian
    from centrex_tlf_amd.lindblad.effective import effective_operator_payload

    rng = np.random.default_rng(seed)
    D = 2 * n * n
    grid = np.sort(rng.uniform(-1.0, 1.0, n_grid))
    grid[0], grid[-1] = -1.0, 1.0
    norm = 1.0 / math.sqrt(density * D)

    def mats(scale):
        mask = rng.random((D, D)) < density
        a, b, c = (rng.normal(size=(D, D)) * mask * norm for _ in range(3))
        return [scale * (a + b * g + 0.3 * c * g * g) - 1.5 * np.eye(D) * (scale > 0.5) for g in grid]

    params = LindbladParameters()
    params.real("delta", 0.3)
    x0 = params.real("x0", -0.9)
    v = params.real("v", 0.35)
    params.bind("field", x0 + v * Time(), finalize=False)
    amp = params.real("Omega_amp", 1.0)
    params.bind("Omega", gaussian(Time(), center=2.5, sigma=1.0, amplitude=amp), finalize=False)
    params._finalize()
    names = list(params.slot_names)
    return {
        "n_states": n, "real_dim": D, "n_grid": n_grid, "field_grid": grid,
        "sparse_combined": effective_operator_payload(mats(1.0), grid, "pchip"),
        "sparse_opt": effective_operator_payload(mats(0.3), grid, "pchip"),
        "sparse_det": effective_operator_payload(mats(0.2), grid, "pchip"),
        "field_coordinate_slot": names.index("field"), "rabi_rate_slot": names.index("Omega"),
        "detuning_slot": names.index("delta"), "operator_interpolation": "pchip", "parameter_graph": params,
    }


def extras_leg(cpu_seconds: float) -> dict:
    """The paths beyond the headline workload, each timed on the GPU with a
    bounded CPU-oracle baseline and a parity sample: fixed-step RK4, a
    population stop event and 101-point save-plan traces on the n=39 64x64
    grid, and the effective-Hamiltonian path on a synthetic n=16 plan."""
    import numpy as np
    import torch

    from centrex_tlf_amd.lindblad import normalize_stop_event, population, prepare_lindblad_problem
    from centrex_tlf_amd.lindblad._driver import build_save_plan
    from centrex_tlf_amd.lindblad.effective import prepare_effective_lindblad_plan, split_real_rho
    from centrex_tlf_amd.lindblad.native import GridAxes
    from oracle.lb_oracle import EffectiveOracle, Oracle, host_threads
    from tests.systems import GAMMA, tlf_n39

    dev = torch.device("cuda")
    cores = host_threads()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    plan = prepared.device_plan()
    oracle = Oracle(prepared)
    names = prepared.slot_names
    slots = [names.index("delta0"), names.index("Omega0")]
    deltas = np.linspace(-5.0 * GAMMA, 5.0 * GAMMA, 64)
    omegas = np.geomspace(0.1 * GAMMA, 10.0 * GAMMA, 64)
    axes = GridAxes([deltas, omegas])
    y0_np = prepared.layout.pack(rho0)
    y0 = torch.as_tensor(y0_np, device=dev)
    support = plan.support_of(y0_np)
    n_traj = axes.count
    rng = np.random.default_rng(20261202 + 5)
    tol = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)

    def timed(fn, reps=3, around_launch=True):
        """Median of ``reps`` timed calls after one warm call; the events
        bracket the launch itself where the entry point records them."""
        fn()
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            if around_launch:
                out = fn(events=(a, b))
            else:
                a.record()
                out = fn()
                b.record()
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b))
        return out, float(np.median(times))

    def sample_size(solve_idx, total):
        """A seeded sample sized to about ``cpu_seconds`` of oracle work on
        all host cores, from the rate of a short probe of 2 x cores."""
        probe = np.sort(rng.choice(total, size=min(total, 2 * cores), replace=False))
        t0 = time.perf_counter()
        solve_idx(probe)
        rate = probe.size / max(time.perf_counter() - t0, 1e-6)
        k = int(min(total, max(2 * cores, rate * cpu_seconds)))
        return np.sort(rng.choice(total, size=k, replace=False))

    def cpu_sample(solve, total=n_traj):
        """Oracle rate on the bounded seeded sample (all host cores)."""
        def solve_idx(idx):
            return solve(np.stack([deltas[idx // 64], omegas[idx % 64]], axis=1).astype(np.complex128))
        idx = sample_size(solve_idx, total)
        t0 = time.perf_counter()
        ref = solve_idx(idx)
        return idx, ref, idx.size / (time.perf_counter() - t0)

    out: dict = {}
    grid_kw = dict(override_slots=slots, grid_axes=axes, support=support)
    # ---- fixed-step RK4, dt = 2 ns (~1000 steps over 20/Gamma)
    dt_fixed = 2e-9
    (pops, stats, _), ms = timed(lambda **e: plan.solve(y0, *t_span, solver="fixed_rk4", dt=dt_fixed,
                                                        output_kind="populations", **grid_kw, **e))
    idx, ref, cpu_rate = cpu_sample(lambda pb: oracle.solve_batch(
        y0_np, t_span, solver="fixed_rk4", dt=dt_fixed, output="populations", output_when="final",
        parameter_slots=["delta0", "Omega0"], parameter_batch=pb, threads=cores))
    err = float(np.max(np.abs(pops.cpu().numpy()[idx] - ref["final"])))
    out["fixed_rk4"] = {"value": n_traj / (ms / 1e3), "unit": "trajectories/s", "kernel_ms": ms,
                        "config": "tlf_n39 64x64 grid, fixed_rk4 dt 2 ns, final populations",
                        "cpu_baseline": {"value": cpu_rate, "unit": "trajectories/s", "cores": cores, "kind": "port",
                                         "sample": f"{len(idx)} seeded grid trajectories, oracle"},
                        "parity": {"trajectories": len(idx), "max_abs_population_error": err, "bar": "1e-12"}}
    log(f"bench: extras fixed_rk4 {out['fixed_rk4']['value']:.0f} traj/s ({ms:.2f} ms), parity {err:.1e}")
    # ---- a population stop event (excited manifold reaches 2 %)
    ev = normalize_stop_event(population([36, 37, 38], threshold=0.02), prepared)
    (pops, stats, _, ev_t), ms = timed(lambda **e: plan.solve(y0, *t_span, output_kind="populations",
                                                              stop_event=ev, **grid_kw, **tol, **e))
    idx, ref, cpu_rate = cpu_sample(lambda pb: oracle.solve_batch(
        y0_np, t_span, output="populations", output_when="final", parameter_slots=["delta0", "Omega0"],
        parameter_batch=pb, stop_event=ev, threads=cores, **tol))
    got_t = ev_t.cpu().numpy()[idx]
    fired = np.isfinite(ref["event_t"])
    terr = float(np.max(np.abs(got_t[fired] / ref["event_t"][fired] - 1.0))) if fired.any() else 0.0
    out["stop_event"] = {"value": n_traj / (ms / 1e3), "unit": "trajectories/s", "kernel_ms": ms,
                         "config": "tlf_n39 64x64 grid, dopri5 rtol 1e-8, population event sum(B) = 0.02",
                         "fired": int(np.isfinite(ev_t.cpu().numpy()).sum()),
                         "cpu_baseline": {"value": cpu_rate, "unit": "trajectories/s", "cores": cores,
                                          "kind": "port", "sample": f"{len(idx)} seeded grid trajectories, oracle"},
                         "parity": {"trajectories": len(idx), "flags_equal": bool(np.array_equal(
                             np.isfinite(got_t), fired)), "max_rel_event_time_error": terr, "bar": "1e-7"}}
    log(f"bench: extras stop_event {out['stop_event']['value']:.0f} traj/s ({ms:.2f} ms), event times {terr:.1e}")
    # ---- save-plan traces: populations at 101 points through the dense output
    sv = np.linspace(t_span[0], t_span[1], 101)[1:]  # with the start: 101 points
    plan_t = build_save_plan(sv, *t_span, True)
    (vals, stats, _, _, counts, _), ms = timed(lambda **e: plan.solve(
        y0, *t_span, output_kind="populations", save_plan=plan_t, trace=True, capacity=plan_t.size, **grid_kw,
        **tol, **e))
    idx, ref, cpu_rate = cpu_sample(lambda pb: oracle.solve_batch(
        y0_np, t_span, output="populations", output_when="saveat", saveat=sv, parameter_slots=["delta0", "Omega0"],
        parameter_batch=pb, threads=cores, **tol))
    err = float(np.max(np.abs(vals.cpu().numpy()[idx] - ref["values"])))
    out["saveat_trace"] = {"value": n_traj / (ms / 1e3), "unit": "trajectories/s", "kernel_ms": ms,
                           "config": "tlf_n39 64x64 grid, dopri5 rtol 1e-8, populations at 101 save points",
                           "cpu_baseline": {"value": cpu_rate, "unit": "trajectories/s", "cores": cores,
                                            "kind": "port", "sample": f"{len(idx)} seeded grid trajectories, oracle"},
                           "parity": {"trajectories": len(idx), "max_abs_population_error": err, "bar": "1e-7"}}
    log(f"bench: extras saveat_trace {out['saveat_trace']['value']:.0f} traj/s ({ms:.2f} ms), parity {err:.1e}")
    # ---- effective-Hamiltonian path: synthetic n=16 plan (512 reals), 4096 trajectories
    eplan = prepare_effective_lindblad_plan(_effective_synthetic_plan(16))
    eor = EffectiveOracle(eplan)
    r0 = np.zeros((16, 16), dtype=complex)
    r0[0, 0] = r0[1, 1] = 0.5
    ey0 = split_real_rho(r0)
    B = 4096
    amps = np.linspace(0.2, 2.0, B)
    eslots = [eplan.slot_names.index("Omega_amp")]
    save = build_save_plan(np.array([5.0]), 0.0, 5.0, False)
    etol = dict(abstol=1e-10, reltol=1e-8, dt=1e-3)
    (evals, estats, _, _, _), ms = timed(lambda: eplan.device.solve(
        ey0, 0.0, 5.0, output_kind="populations", override_slots=eslots, override_values=amps[:, None], n_traj=B,
        save_plan=save, save_start=False, capacity=1, **etol), around_launch=False)
    def eff_solve(idx):
        return eor.solve_batch(ey0, (0.0, 5.0), output="populations", output_when="final",
                               parameter_slots=["Omega_amp"], parameter_batch=amps[idx, None].astype(complex),
                               threads=cores, **etol)
    eidx = sample_size(eff_solve, B)
    t0 = time.perf_counter()
    eref = eff_solve(eidx)
    ecpu = eidx.size / (time.perf_counter() - t0)
    # the synthetic operators damp the state (trace ~1e-3 at t1): error relative to the largest value
    err = float(np.max(np.abs(evals.cpu().numpy()[eidx, 0] - eref["final"])) / np.max(np.abs(eref["final"])))
    out["effective"] = {"value": B / (ms / 1e3), "unit": "trajectories/s", "kernel_ms": ms,
                        "timing": "events around the call (host argument set-up included)",
                        "config": "synthetic effective plan n=16 (512 reals, ~5.5k stored entries per operator, "
                                  "6 field points, pchip), 4096 Rabi amplitudes, dopri5 rtol 1e-8, populations at t1",
                        "accepted_steps_mean": float(estats.cpu().numpy()[:, 0].mean()),
                        "cpu_baseline": {"value": ecpu, "unit": "trajectories/s", "cores": cores, "kind": "port",
                                         "sample": f"{len(eidx)} seeded trajectories, oracle"},
                        "parity": {"trajectories": len(eidx), "max_rel_population_error": err, "bar": "1e-6"}}
    log(f"bench: extras effective {out['effective']['value']:.0f} traj/s ({ms:.2f} ms), parity {err:.1e}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--solver", default="dopri5", choices=["dopri5", "tsit5"])
    ap.add_argument("--threads-per-traj", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="kernel variant (0 = plan default)")
    ap.add_argument("--workgroups", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--plain-sample", type=int, default=128)
    ap.add_argument("--profile", action="store_true", help="GPU steps only (for rocprofv3)")
    ap.add_argument("--full-plan", action="store_true", help="integrate every element (no restriction)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="launch eagerly instead of replaying a HIP graph of one step")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the adjoint-gradient leg")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint-gradient leg")
    ap.add_argument("--no-uneven", dest="uneven", action="store_false", help="skip the uneven-batch leg")
    ap.add_argument("--fit", action="store_true", help="run only the batched fit leg")
    ap.add_argument("--fit-trajectories", type=int, default=2048, help="trajectories per system in the fit")
    ap.add_argument("--fit-iterations", type=int, default=50)
    ap.add_argument("--fit-time-scale", type=float, default=1.0, help="scale of the transit times in the fit")
    ap.add_argument("--fit-eager", dest="fit_graph", action="store_false",
                    help="run the fit's iterations eagerly (default: replayed from a HIP graph)")
    ap.add_argument("--uneven-only", action="store_true", help="run only the uneven-batch leg")
    ap.add_argument("--no-specialize", dest="specialize", action="store_false",
                    help="run the precompiled kernels (expression machine) instead of the plan-specialised ones")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="skip the fixed-step, stop-event, trace and effective-path legs")
    ap.add_argument("--extras-only", action="store_true", help="run only those legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; --gpus must be 1 (on several GPUs each "
                                   "would solve its own contiguous range of trajectories)"}))
        return 2

    import numpy as np
    import torch

    from centrex_tlf_amd.lindblad import prepare_lindblad_problem
    from centrex_tlf_amd.lindblad.native import GridAxes
    from tests.systems import GAMMA, tlf_n39

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 1
    dev = torch.device("cuda", 0)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": 5}))
        return 0
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(args.steps, args.warmup)))
        return 0
    if args.fit:
        print(json.dumps(fit_leg(args.fit_trajectories, args.fit_iterations, args.fit_time_scale, args.fit_graph)))
        return 0
    if args.uneven_only:
        print(json.dumps(uneven_leg(min(args.steps, 3), 1, args.cpu_seconds)))
        return 0
    if args.extras_only:
        print(json.dumps(extras_leg(min(args.cpu_seconds, 4.0))))
        return 0
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    prepared.specialize = args.specialize
    plan = prepared.device_plan()
    info = plan.info
    slot_names = prepared.slot_names
    slots = [slot_names.index("delta0"), slot_names.index("Omega0")]
    deltas = np.linspace(-5.0 * GAMMA, 5.0 * GAMMA, 64)
    omegas = np.geomspace(0.1 * GAMMA, 10.0 * GAMMA, 64)
    axes = GridAxes([deltas, omegas])
    y0 = torch.as_tensor(prepared.layout.pack(rho0), device=dev)
    tol = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
    n_traj = axes.count
    support = plan.support_of(prepared.layout.pack(rho0))
    run_kw = dict(restrict_support=False) if args.full_plan else dict(support=support)
    rinfo = info if args.full_plan else plan.support_info(support)

    def step(**kw):
        return plan.solve(y0, t_span[0], t_span[1], solver=args.solver, output_kind="populations",
                          weights=weights, with_integral=True, override_slots=slots, grid_axes=axes,
                          threads_per_traj=args.threads_per_traj, workgroups=args.workgroups,
                          variant=args.variant, **run_kw, **tol, **kw)

    log(f"bench: plan {info}")
    log(f"bench: solved tables {rinfo}")
    for i in range(args.warmup):
        out = step()
        torch.cuda.synchronize()
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    graph = None
    if args.graph:
        if args.warmup == 0:  # a specialised plan compiles on its first call, not inside a capture
            out = step()
            torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            out = step()
        torch.cuda.synchronize()
        log("bench: captured one step in a HIP graph")
    torch.cuda.synchronize()
    t_begin = time.perf_counter()
    for i in range(args.steps):
        if graph is not None:
            starts[i].record()
            graph.replay()
            ends[i].record()
        else:
            out = step(events=(starts[i], ends[i]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_begin
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    pops, stats, fail_t, photons = out
    stats_np = stats.cpu().numpy()
    if np.any(stats_np[:, 3] != 0):
        bad = int(np.argmax(stats_np[:, 3] != 0))
        log(f"bench: trajectory {bad} failed with status {stats_np[bad, 3]}")
        return 1
    log(f"bench: {args.steps} steps in {elapsed:.3f} s")
    flops = algorithmic_flops(stats_np, info["flops_per_rhs"], info["dim"], args.solver)
    executed = algorithmic_flops(stats_np, rinfo["flops_per_rhs"], rinfo["dim"], args.solver)
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    achieved_tflops = flops / avg_kernel_s / 1e12
    lds_rhs = lds_bytes_per_rhs(prepared)
    lds_bytes = float(stats_np[:, 2].astype(np.float64).sum()) * lds_rhs
    traffic = None
    if TRAFFIC_FILE.exists():
        traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")
    result = {
        "metric": "Lindblad trajectories/s (batched Dopri5, n=39, rtol 1e-8); adjoint grads/s",
        "value": n_traj * args.steps / elapsed,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (TlF n=39 system from the committed fixture; 64x64 detuning x Rabi grid)",
        "config": {
            "workload": "tlf_n39_grid_64x64",
            "n_states": 39,
            "trajectories": n_traj,
            "solver": args.solver,
            "rtol": 1e-8,
            "atol": 1e-10,
            "outputs": "final populations + photon count (one launch)",
            "parallelism": "single_gpu",
            "integrated_elements": f"{rinfo['upper_len']} of {info['upper_len']} upper-triangle elements",
            "kernel_variant": args.variant or plan.launch_info(support=None if args.full_plan else support)["variant"],
            "hip_graph": bool(args.graph),
            "plan_specialized": bool(plan.specialized),
            "threads_per_traj": args.threads_per_traj or rinfo["threads_per_traj"],
        },
        "roofline": {
            "bound": "fp64_valu",
            "latency_bound": True,
            "compute": "FP64 vector ALU (no MFMA: the sparse RHS has no dense tile); the MI355X FP64 vector "
                       "and matrix peaks are both 78.6 TFLOP/s.  Each trajectory's step chain is serial, so the "
                       "launch time is the longest trajectory's attempts times the per-attempt latency",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel_ms": float(np.mean(kernel_ms)),
            "flops_per_launch": flops,
            "flops_definition": "sum over trajectories of rhs_calls*F_rhs + attempted*dim*F_rk, "
                                f"F_rhs={info['flops_per_rhs']} (full plan), dim={info['dim']}, "
                                f"F_rk={F_RK[args.solver]}",
            "executed_flops_per_launch": executed,
            "executed_frac": executed / avg_kernel_s / 1e12 / FP64_PEAK_TFLOPS,
            "lds_cobound": {
                "achieved": lds_bytes / avg_kernel_s / 1e12,
                "peak": LDS_PEAK_TBS,
                "unit": "TB/s",
                "frac": lds_bytes / avg_kernel_s / 1e12 / LDS_PEAK_TBS,
                "bytes_per_launch": lds_bytes,
                "definition": "sum over trajectories of rhs_calls * (sum over terms of 8 bytes per real input, 16 "
                              f"per complex, + 8*dim), full plan: {lds_rhs} bytes per RHS; peak 256 CUs x 128 B/clk "
                              "x 2.4 GHz",
            },
        },
        "totals": {
            "accepted_steps": int(stats_np[:, 0].sum()),
            "rejected_steps": int(stats_np[:, 1].sum()),
            "rhs_calls": int(stats_np[:, 2].sum()),
            "steps_min_max": [int(stats_np[:, 0].min()), int(stats_np[:, 0].max())],
        },
    }
    if args.profile:
        print(json.dumps(result))
        return 0
    if not args.full_plan:
        # ---- the same grid on the full plan: every one of the 780 upper
        # elements integrated, the work the reference's kernels do
        def full_step(events=None):
            return plan.solve(y0, t_span[0], t_span[1], solver=args.solver, output_kind="populations",
                              weights=weights, with_integral=True, override_slots=slots, grid_axes=axes,
                              restrict_support=False, events=events, **tol)

        full_out = full_step()
        torch.cuda.synchronize()
        fev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        t_f = time.perf_counter()
        for e in fev:
            full_out = full_step(events=e)
        torch.cuda.synchronize()
        full_s = (time.perf_counter() - t_f) / len(fev)
        full_kms = float(np.mean([a.elapsed_time(b) for a, b in fev]))
        full_stats = full_out[1].cpu().numpy()
        full_flops = algorithmic_flops(full_stats, info["flops_per_rhs"], info["dim"], args.solver)
        result["full_plan"] = {
            "value": n_traj / full_s,
            "unit": "trajectories/s",
            "kernel_ms": full_kms,
            "frac": full_flops / (full_kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "achieved_tflops": full_flops / (full_kms * 1e-3) / 1e12,
            "max_abs_population_difference_vs_restricted": float((full_out[0] - pops).abs().max().item()),
            "accepted_steps_equal": bool(np.array_equal(full_stats[:, 0], stats_np[:, 0])),
        }
        log(f"bench: full plan {n_traj / full_s:.0f} traj/s, kernel {full_kms:.2f} ms")

    rng = np.random.default_rng(20261202 + 1)
    # ---- plain GPU path: one launch per trajectory on a seeded sample
    sample = np.sort(rng.choice(n_traj, size=min(args.plain_sample, n_traj), replace=False))
    pb = np.stack([deltas[sample // 64], omegas[sample % 64]], axis=1).astype(np.complex128)
    pb_dev = torch.as_tensor(pb, device=dev)
    plan.solve(y0, *t_span, solver=args.solver, output_kind="populations", weights=weights,
               with_integral=True, override_slots=slots, override_values=pb_dev[:2], mode="plain",
               **run_kw, **tol)
    torch.cuda.synchronize()
    t_p = time.perf_counter()
    p_pops, p_stats, _, p_phot = plan.solve(y0, *t_span, solver=args.solver, output_kind="populations",
                                            weights=weights, with_integral=True, override_slots=slots,
                                            override_values=pb_dev, mode="plain", **run_kw, **tol)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t_p
    plain_rate = len(sample) / plain_s
    # the plain path must reproduce the batched results bit for bit
    plain_equal = bool(torch.equal(p_pops, pops[torch.as_tensor(sample, device=dev)]) and
                       torch.equal(p_phot, photons[torch.as_tensor(sample, device=dev)]))
    log(f"bench: plain path {plain_rate:.1f} traj/s on {len(sample)} trajectories")
    result["plain_gpu"] = {"value": plain_rate, "unit": "trajectories/s", "sample": len(sample),
                           "bitwise_equal_to_batched": plain_equal}
    result["speedup_vs_plain"] = result["value"] / plain_rate

    # ---- CPU baseline: the oracle on all host cores, bounded seeded sample
    from oracle.lb_oracle import Oracle, host_threads

    oracle = Oracle(prepared)
    cores = host_threads()
    y0_np = prepared.layout.pack(rho0)
    probe = np.sort(rng.choice(n_traj, size=cores, replace=False))
    pbp = np.stack([deltas[probe // 64], omegas[probe % 64]], axis=1).astype(np.complex128)
    t_c = time.perf_counter()
    oracle.solve_batch(y0_np, t_span, solver=args.solver, output="populations", output_when="final",
                       parameter_slots=["delta0", "Omega0"], parameter_batch=pbp, threads=cores, **tol)
    probe_s = time.perf_counter() - t_c
    n_cpu = int(max(cores, min(n_traj, cores * math.floor(args.cpu_seconds / max(probe_s, 1e-3)))))
    cpu_idx = np.sort(rng.choice(n_traj, size=n_cpu, replace=False))
    pbc = np.stack([deltas[cpu_idx // 64], omegas[cpu_idx % 64]], axis=1).astype(np.complex128)
    log(f"bench: cpu baseline on {n_cpu} trajectories x {cores} threads")
    t_c = time.perf_counter()
    ref = oracle.solve_batch(y0_np, t_span, solver=args.solver, output="populations", output_when="final",
                             parameter_slots=["delta0", "Omega0"], parameter_batch=pbc, threads=cores, **tol)
    cpu_s = time.perf_counter() - t_c
    cpu_rate = n_cpu / cpu_s
    ref_int = oracle.solve_batch(y0_np, t_span, solver=args.solver, output="photon_integral",
                                 output_when="final", integral_weights=weights,
                                 parameter_slots=["delta0", "Omega0"], parameter_batch=pbc,
                                 threads=cores, **tol)
    result["cpu_baseline"] = {
        "value": cpu_rate,
        "unit": "trajectories/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n_cpu} seeded trajectories of the same 64x64 grid (final populations), "
                  f"oracle C restatement of the reference Rust path, {cores} threads, {cpu_s:.1f} s",
    }
    result["speedup_vs_cpu"] = result["value"] / cpu_rate
    # one host thread, on a smaller seeded sample of the same grid
    n_one = int(max(4, min(n_cpu, cpu_rate / cores * args.cpu_seconds / 4)))
    pb1 = pbc[:n_one]
    t_c = time.perf_counter()
    oracle.solve_batch(y0_np, t_span, solver=args.solver, output="populations", output_when="final",
                       parameter_slots=["delta0", "Omega0"], parameter_batch=pb1, threads=1, **tol)
    one_s = time.perf_counter() - t_c
    result["cpu_baseline_one_thread"] = {"value": n_one / one_s, "unit": "trajectories/s", "cores": 1,
                                         "kind": "port", "sample": f"{n_one} of the seeded trajectories above"}
    # ---- the single trajectory of the CPU configuration (delta0 = 0, Omega0 =
    # Gamma): the reference needs two solves for populations and the photon
    # count; the GPU gives both from one launch of the per-trajectory path
    one = np.asarray([[0.0, GAMMA]], dtype=np.complex128)
    t_c = time.perf_counter()
    c1_pop = oracle.solve_batch(y0_np, t_span, solver=args.solver, output="populations", output_when="final",
                                parameter_slots=["delta0", "Omega0"], parameter_batch=one, threads=1, **tol)
    c1_int = oracle.solve_batch(y0_np, t_span, solver=args.solver, output="photon_integral", output_when="final",
                                integral_weights=weights, parameter_slots=["delta0", "Omega0"],
                                parameter_batch=one, threads=1, **tol)
    c1_cpu = time.perf_counter() - t_c
    one_dev = torch.as_tensor(one, device=dev)
    c1_kw = dict(solver=args.solver, output_kind="populations", weights=weights, with_integral=True,
                 override_slots=slots, override_values=one_dev, mode="plain", **run_kw, **tol)
    plan.solve(y0, *t_span, **c1_kw)
    torch.cuda.synchronize()
    t_c = time.perf_counter()
    reps = 20
    for _ in range(reps):
        c1_out = plan.solve(y0, *t_span, **c1_kw)
    torch.cuda.synchronize()
    c1_gpu = (time.perf_counter() - t_c) / reps
    result["single_trajectory"] = {
        "config": "tlf_n39, delta0 = 0, Omega0 = Gamma, Dopri5 rtol 1e-8: final populations and photon count",
        "cpu_seconds_one_thread": c1_cpu,
        "gpu_seconds_plain_path": c1_gpu,
        "accepted_steps": int(c1_pop["accepted"][0]),
        "photon_count": float(c1_int["final"][0, 0]),
        "max_abs_population_error": float(np.abs(c1_out[0].cpu().numpy()[0] - c1_pop["final"][0]).max()),
        "photon_error": float(abs(c1_out[3].cpu().numpy()[0] - c1_int["final"][0, 0])),
    }
    # ---- parity of the same sample: populations, photon counts, accepted steps
    idx_dev = torch.as_tensor(cpu_idx, device=dev)
    gpu_pops = pops[idx_dev].cpu().numpy()
    gpu_phot = photons[idx_dev].cpu().numpy()
    gpu_steps = stats_np[cpu_idx, 0]
    result["parity"] = {
        "trajectories": n_cpu,
        "max_abs_population_error": float(np.max(np.abs(gpu_pops - ref["final"]))),
        "max_abs_photon_error": float(np.max(np.abs(gpu_phot - ref_int["final"][:, 0]))),
        "max_trace_error": float(np.max(np.abs(pops.sum(dim=1).cpu().numpy() - 1.0))),
        "accepted_steps_equal": bool(np.array_equal(gpu_steps, ref["accepted"])),
        "bar": "1e-7 absolute, trace 1e-10",
    }
    if args.adjoint:
        result["adjoint"] = adjoint_leg(args.steps, args.warmup)
    if args.uneven:
        result["uneven"] = uneven_leg(min(args.steps, 3), 1, args.cpu_seconds)
    if args.extras:
        result["extras"] = extras_leg(min(args.cpu_seconds, 4.0))
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
