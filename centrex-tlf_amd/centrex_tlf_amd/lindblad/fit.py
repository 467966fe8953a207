"""Batched fit of per-trajectory parameters to photon-count targets.

Each group is one prepared system with its own trajectories (initial state,
end times, fixed per-trajectory parameters such as a beam width) and
photon-count targets.  Every iteration runs, per group, the forward solve
that records the adjoint tape and the discrete-adjoint kernel
(``adjoint.DifferentiablePhotonCount``), then one Adam step on all
parameters.  The loss is sum_b (N_b - target_b)^2 over every trajectory of
every group; the parameters are optimised in units of ``scales`` (e.g.
Gamma for detunings and Rabi rates, 1 for angles).

The reference has no fitting or gradients; the photon count is its
``photon_integral`` output (``rust/src/ode/output.rs:253-335``).
"""

from __future__ import annotations

import time
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .adjoint import DifferentiablePhotonCount

__all__ = ["FitGroup", "FitResult", "fit_photon_counts"]


@dataclass
class FitGroup:
    """Trajectories of one system: ``rho0`` (n, n) or packed (dim,) shared by
    the group; ``t_end`` (B,) end times; ``targets`` (B,) photon counts;
    ``initial`` (B, P) starting values of the fitted parameters; ``fixed``
    maps other base parameters to per-trajectory values (B,)."""

    prepared: Any
    rho0: np.ndarray
    t_end: np.ndarray
    targets: np.ndarray
    initial: np.ndarray
    integral_weights: Sequence[tuple[int, float]]
    fixed: dict[str, np.ndarray] = field(default_factory=dict)


@dataclass
class FitResult:
    parameters: list[np.ndarray]     # per group (B, P), physical units
    loss: list[float]                # per iteration, before its step
    counts: list[np.ndarray]         # per group (B,), at the returned parameters
    seconds: float                   # wall time of the iterations
    gradients_per_second: float      # trajectory gradients (all P parameters) per second
    tape_bytes: int                  # peak adjoint tape memory of one iteration (all groups)
    forward_seconds: float = 0.0     # forward solves with tape, all iterations
    adjoint_seconds: float = 0.0     # adjoint kernels, all iterations


def fit_photon_counts(groups: Sequence[FitGroup], parameters: Sequence[str], *, scales: Sequence[float],
                      iterations: int = 50, lr: float = 0.05, solver: str = "dopri5", abstol: float = 1e-10,
                      reltol: float = 1e-8, dt: float = 1e-8, maxiters: int = 100_000, ckpt_every: int = 16,
                      tape_capacity: int = 1024, callback=None) -> FitResult:
    """Adam on the parameters of every trajectory of every group (see module docstring)."""
    import torch

    if not groups:
        raise ValueError("a fit needs at least one group")
    scale = torch.as_tensor(np.asarray(scales, dtype=np.float64), device="cuda")
    if scale.numel() != len(parameters):
        raise ValueError("one scale per fitted parameter")
    engines, y0s, targets, us = [], [], [], []
    for g, grp in enumerate(groups):
        t_end = np.ascontiguousarray(grp.t_end, dtype=np.float64).reshape(-1)
        init = np.asarray(grp.initial, dtype=np.float64)
        if init.shape != (t_end.size, len(parameters)):
            raise ValueError(f"group {g}: initial must have shape ({t_end.size}, {len(parameters)})")
        eng = DifferentiablePhotonCount(grp.prepared, parameters, grp.integral_weights, (0.0, t_end),
                                        solver=solver, abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters,
                                        ckpt_every=ckpt_every, tape_capacity=tape_capacity, fixed=grp.fixed)
        r = np.asarray(grp.rho0)
        y0 = grp.prepared.layout.pack(r) if r.ndim == 2 else np.ascontiguousarray(r, dtype=np.float64)
        engines.append(eng)
        y0s.append(torch.as_tensor(y0, device="cuda"))
        targets.append(torch.as_tensor(np.asarray(grp.targets, dtype=np.float64), device="cuda"))
        us.append((torch.as_tensor(init, device="cuda") / scale).requires_grad_(True))
    opt = torch.optim.Adam(us, lr=lr)
    losses: list[float] = []
    t_fwd = t_adj = 0.0
    tape_peak = 0
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for it in range(iterations):
        opt.zero_grad()
        total = 0.0
        tape_bytes = 0
        for eng, y0, tgt, u in zip(engines, y0s, targets, us):
            p = (u * scale).detach()
            t0 = time.perf_counter()
            counts, tape = eng.forward(p, y0)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            dIdp = eng.gradient(p, y0, tape)
            torch.cuda.synchronize()
            t_fwd += t1 - t0
            t_adj += time.perf_counter() - t1
            tape_bytes += tape.nbytes
            resid = counts - tgt
            total += float((resid * resid).sum().item())
            # d loss / d u = 2 (N - target) dN/dp * scale
            u.grad = (2.0 * resid[:, None] * dIdp * scale).detach()
            del tape
        tape_peak = max(tape_peak, tape_bytes)
        losses.append(total)
        if callback is not None:
            callback(it, total)
        opt.step()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t_start
    params = [(u * scale).detach() for u in us]
    counts = [eng.forward(p, y0)[0].cpu().numpy() for eng, p, y0 in zip(engines, params, y0s)]
    n_traj = sum(len(g.targets) for g in groups)
    return FitResult([p.cpu().numpy() for p in params], losses, counts, seconds,
                     n_traj * iterations / seconds, tape_peak, t_fwd, t_adj)
