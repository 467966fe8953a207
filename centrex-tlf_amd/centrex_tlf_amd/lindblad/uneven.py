"""Uneven batches: several OBE systems and per-trajectory end times in one launch.

The reference solves one plan with one ``t_span`` per call
(``centrex_tlf/lindblad/batch.py:209``; ``rust/src/lindblad/ode_batch.rs:176-351``).
Here a batch is a list of groups, each with its own prepared problem (level
count n, Hamiltonian, decay), initial state, parameter rows and end times.
Every group runs its persistent work-queue kernel; the groups run
concurrently on streams forked from the caller's stream (``lb_solve_group``),
so a HIP graph captured around :func:`solve_uneven_batch` replays the whole
batch as one launch.  Within a group the queue hands out trajectories longest
first (by end time), which shortens the tail of uneven step counts.

The launch parameters -- kernel variant (slots per thread, gather rows in
registers, workgroup size) per group and each group's share of the device
(persistent workgroups) -- are chosen by :func:`autotune`, which times the
candidates on the batch itself and caches the choice.  All variants round
identically, so tuning never changes a result.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _driver as D
from .native import NativeError, kernel_variants, solve_group

__all__ = ["UnevenGroup", "UnevenGroupResult", "UnevenTuning", "autotune", "solve_uneven_batch"]


@dataclass
class UnevenGroup:
    """Trajectories of one system: ``rho0`` is (n, n) shared, (B, n, n), or
    packed (dim,) / (B, dim); ``t_end`` (B,) end times; ``parameter_batch``
    (B, P) values of ``parameter_slots``; ``integral_weights`` as the
    reference's photon-count output (``[(state, weight), ...]``) or None."""

    prepared: Any
    rho0: np.ndarray
    t_end: np.ndarray
    parameter_slots: Sequence[Any] | None = None
    parameter_batch: np.ndarray | None = None
    integral_weights: Sequence[tuple[int, float]] | None = None


@dataclass
class UnevenGroupResult:
    """Final populations (B, n), photon counts (B,) when weights were given,
    and per-trajectory stats (B, 4): accepted, rejected, rhs calls, status."""

    populations: np.ndarray
    photons: np.ndarray | None
    stats: np.ndarray


@dataclass
class UnevenTuning:
    """Launch parameters of an uneven batch: per group (variant, threads per
    trajectory, persistent workgroups; 0 = defaults / fill the device)."""

    variants: list[int]
    threads: list[int]
    workgroups: list[int]
    mode: str = "default"  # "default", "sequential" or "concurrent"
    measured_ms: float = float("nan")
    trials: list[dict[str, Any]] = field(default_factory=list)


_CACHE: dict[tuple, UnevenTuning] = {}


def _pack_y0(prepared: Any, rho0: np.ndarray, count: int) -> np.ndarray:
    r = np.asarray(rho0)
    n, dim = prepared.layout.n, prepared.layout.packed_len
    if r.ndim == 1 and r.size == dim:
        return np.ascontiguousarray(r, dtype=np.float64)
    if r.ndim == 2 and r.shape == (n, n):
        return np.ascontiguousarray(prepared.layout.pack(r), dtype=np.float64)
    if r.ndim == 2 and r.shape == (count, dim):
        return np.ascontiguousarray(r, dtype=np.float64)
    if r.ndim == 3 and r.shape == (count, n, n):
        return np.ascontiguousarray(np.stack([prepared.layout.pack(x) for x in r]), dtype=np.float64)
    raise ValueError(f"rho0 must be (n, n), (B, n, n), (dim,) or (B, dim) with n = {n}, B = {count}")


def _group_calls(groups: Sequence[UnevenGroup], *, t0: float, solver: str, abstol: float, reltol: float,
                 dt: float, maxiters: int, save_start: bool, longest_first: bool):
    if not groups:
        raise ValueError("an uneven batch needs at least one group")
    if solver not in D.GPU_SOLVERS:
        raise NotImplementedError(f"solver {solver!r} is not available on the GPU")
    calls = []
    for g, grp in enumerate(groups):
        t_end = np.ascontiguousarray(grp.t_end, dtype=np.float64).reshape(-1)
        count = t_end.size
        if count == 0:
            raise ValueError(f"group {g}: no trajectories")
        if not np.all(np.isfinite(t_end)):
            raise ValueError(f"group {g}: end times must be finite")
        y0 = _pack_y0(grp.prepared, grp.rho0, count)
        kw: dict[str, Any] = dict(y0=y0, t0=float(t0), t1=t_end, solver=solver, abstol=float(abstol),
                                  reltol=float(reltol), dt=float(dt), maxiters=int(maxiters),
                                  save_start=bool(save_start), output_kind="populations")
        if grp.integral_weights is not None:
            kw.update(weights=[(int(i), float(w)) for i, w in grp.integral_weights], with_integral=True)
        slots = D.parameter_slot_indices(grp.prepared, grp.parameter_slots)
        if grp.parameter_batch is not None:
            vals = np.ascontiguousarray(grp.parameter_batch, dtype=np.complex128)
            if vals.shape != (count, len(slots)):
                raise ValueError(f"group {g}: parameter_batch must have shape ({count}, {len(slots)})")
            kw.update(override_slots=slots, override_values=vals)
        elif slots:
            raise ValueError(f"group {g}: parameter_slots were provided without parameter_batch")
        else:
            kw.update(n_traj=count)
            if y0.ndim == 2:
                kw.pop("n_traj")
        if longest_first:
            kw.update(queue_order=np.argsort(-t_end, kind="stable").astype(np.int32))
        # the restricted tables follow the nonzero entries of y0 (computed on
        # the host here, so a solve under graph capture reads nothing back)
        kw.update(support=((y0 != 0) if y0.ndim == 1 else (y0 != 0).any(axis=0)).astype(np.uint8))
        calls.append((grp.prepared, kw))
    return [(prepared.device_plan(), kw) for prepared, kw in calls]


def _with_tuning(calls, tuning: UnevenTuning | None):
    if tuning is None:
        return calls
    out = []
    for g, (plan, kw) in enumerate(calls):
        k = dict(kw)
        k.update(variant=int(tuning.variants[g]), threads_per_traj=int(tuning.threads[g]),
                 workgroups=int(tuning.workgroups[g]))
        out.append((plan, k))
    return out


def _time_ms(calls, repeats: int) -> float:
    import torch

    best = float("inf")
    for _ in range(repeats):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        solve_group(calls, events=(a, b))
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def _candidates(plan, support) -> list[dict[str, int]]:
    """Compiled variants that can run this plan (1-based ids, plan default first)."""
    out, seen = [], set()
    for v in [0] + [k["variant"] for k in kernel_variants()]:
        try:
            info = plan.launch_info(support=support, variant=v)
        except NativeError:
            continue
        key = (info["variant"], info["threads_per_traj"])
        if key not in seen:
            seen.add(key)
            out.append(info)
    return out


def autotune(groups: Sequence[UnevenGroup], *, t0: float = 0.0, solver: str = "dopri5", abstol: float = 1e-10,
             reltol: float = 1e-8, dt: float = 1e-8, maxiters: int = 100_000, save_start: bool = True,
             longest_first: bool = True, repeats: int = 2, cache: bool = True) -> UnevenTuning:
    """Pick variants and device shares for this batch by timing them on it.

    1. Each group alone on the whole device, every compatible variant: the
       fastest variant and its time T_g.
    2. The groups together: sequential (each fills the device in turn) and
       concurrent splits of the CUs in proportion to T_g (and to T_g**0.5 and
       equal shares), workgroups_g = share_g * CUs * resident workgroups per CU.
    The fastest configuration is returned (and cached for batches of the same
    plans, sizes, solver and tolerances).
    """
    calls = _group_calls(groups, t0=t0, solver=solver, abstol=abstol, reltol=reltol, dt=dt,
                         maxiters=maxiters, save_start=save_start, longest_first=longest_first)
    key = (tuple(id(p) for p, _ in calls), tuple(int(np.size(kw["t1"])) for _, kw in calls), solver,
           float(abstol), float(reltol), float(dt), bool(longest_first))
    if cache and key in _CACHE:
        return _CACHE[key]
    trials: list[dict[str, Any]] = []
    best_var: list[dict[str, int]] = []
    solo_ms: list[float] = []
    for g, (plan, kw) in enumerate(calls):
        support = kw["support"]
        best = None
        # candidates are ranked on the precompiled kernels: a specialised plan
        # would compile every candidate (seconds each); the chosen ones
        # compile once below
        specialized = getattr(plan, "specialized", False)
        if specialized:
            plan.specialize(False)
        try:
            for info in _candidates(plan, support):
                k = dict(kw, variant=info["variant"], threads_per_traj=info["threads_per_traj"], workgroups=0)
                ms = _time_ms([(plan, k)], repeats)
                trials.append({"group": g, "variant": info["variant"], "threads": info["threads_per_traj"],
                               "solo_ms": ms})
                if best is None or ms < best[0]:
                    best = (ms, info)
        finally:
            if specialized:
                plan.specialize(True)
        if best is None:
            raise NativeError(f"group {g}: no kernel variant can run this plan")
        solo_ms.append(best[0])
        best_var.append(best[1])
    variants = [b["variant"] for b in best_var]
    threads = [b["threads_per_traj"] for b in best_var]
    if any(getattr(p, "specialized", False) for p, _ in calls):  # compile the chosen variants, untimed
        import torch

        solve_group(_with_tuning(calls, UnevenTuning(variants, threads, [0] * len(calls), mode="sequential")))
        torch.cuda.synchronize()
    configs = [("sequential", [0] * len(calls))]
    if len(calls) > 1:
        cus = best_var[0]["cus"]
        for power in (1.0, 0.5, 0.0):
            w = np.asarray(solo_ms) ** power
            share = w / w.sum()
            wg = [max(1, int(round(s * cus))) * max(1, b["per_cu"]) for s, b in zip(share, best_var)]
            configs.append((f"concurrent (share ~ T^{power:g})", wg))
    best_cfg = None
    for name, wg in configs:
        tun = UnevenTuning(variants, threads, wg, mode=name.split(" ")[0])
        ms = _time_ms(_with_tuning(calls, tun), repeats)
        trials.append({"config": name, "workgroups": wg, "ms": ms})
        if best_cfg is None or ms < best_cfg.measured_ms:
            tun.measured_ms = ms
            best_cfg = tun
    best_cfg.trials = trials
    if cache:
        _CACHE[key] = best_cfg
    return best_cfg


def solve_uneven_batch(groups: Sequence[UnevenGroup], *, t0: float = 0.0, solver: str = "dopri5",
                       abstol: float = 1e-10, reltol: float = 1e-8, dt: float = 1e-8, maxiters: int = 100_000,
                       save_start: bool = True, tuning: UnevenTuning | str | None = "auto",
                       longest_first: bool = True, stream=None) -> list[UnevenGroupResult]:
    """Integrate every group from t0 to its trajectories' end times in one call.

    ``tuning``: "auto" (the cached :func:`autotune` choice), an
    :class:`UnevenTuning`, or None (plan defaults; each group then fills the
    device in turn).  Final populations and, for groups with integral weights,
    the photon count (trapezoid over t0 and every accepted step,
    ``rust/src/ode/output.rs:253-335``).  A failed trajectory raises
    ``ValueError("group g, trajectory b: <reference message>")``.
    """
    if isinstance(tuning, str):
        if tuning != "auto":
            raise ValueError("tuning must be 'auto', an UnevenTuning or None")
        tuning = autotune(groups, t0=t0, solver=solver, abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters,
                          save_start=save_start, longest_first=longest_first)
    calls = _with_tuning(_group_calls(groups, t0=t0, solver=solver, abstol=abstol, reltol=reltol, dt=dt,
                                      maxiters=maxiters, save_start=save_start, longest_first=longest_first),
                         tuning)
    outs = solve_group(calls, stream=stream)
    results = []
    for g, (out, grp) in enumerate(zip(outs, groups)):
        values, stats, fail_t = out[0].cpu().numpy(), out[1].cpu().numpy(), out[2].cpu().numpy()
        bad = np.flatnonzero(stats[:, 3] != 0)
        if bad.size:
            i = int(bad[0])
            msg = D.trajectory_message(int(stats[i, 3]), float(fail_t[i]), maxiters=maxiters, n_plan=0,
                                       solver=solver)
            raise ValueError(f"group {g}, trajectory {i}: {msg}")
        photons = out[3].cpu().numpy() if grp.integral_weights is not None else None
        results.append(UnevenGroupResult(values, photons, stats))
    return results
