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
                      tape_capacity: int = 1024, callback=None, graph: bool = False,
                      on_replay=None) -> FitResult:
    """Adam on the parameters of every trajectory of every group (see module docstring).

    With ``graph`` every iteration after the first is one replay of a HIP
    graph that holds all groups' forward solves, adjoints, the loss and the
    Adam step (:func:`_fit_graph`); ``callback`` then runs after the fit, and
    ``on_replay(iteration)`` right after each replay is queued (it may
    synchronise, e.g. to report progress).
    """
    import torch

    if graph:
        return _fit_graph(groups, parameters, scales=scales, iterations=iterations, lr=lr, solver=solver,
                          abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters, ckpt_every=ckpt_every,
                          tape_capacity=tape_capacity, callback=callback, on_replay=on_replay)

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


_ADAM = (0.9, 0.999, 1e-8)  # torch.optim.Adam's defaults


def _engines(groups, parameters, **kw):
    import torch

    out = []
    for g, grp in enumerate(groups):
        t_end = np.ascontiguousarray(grp.t_end, dtype=np.float64).reshape(-1)
        init = np.asarray(grp.initial, dtype=np.float64)
        if init.shape != (t_end.size, len(parameters)):
            raise ValueError(f"group {g}: initial must have shape ({t_end.size}, {len(parameters)})")
        eng = DifferentiablePhotonCount(grp.prepared, parameters, grp.integral_weights, (0.0, t_end),
                                        fixed=grp.fixed, **kw)
        r = np.asarray(grp.rho0)
        y0 = grp.prepared.layout.pack(r) if r.ndim == 2 else np.ascontiguousarray(r, dtype=np.float64)
        tgt = torch.as_tensor(np.asarray(grp.targets, dtype=np.float64), device="cuda")
        out.append((eng, y0, tgt, torch.as_tensor(init, device="cuda")))
    return out


def _fit_graph(groups, parameters, *, scales, iterations, lr, solver, abstol, reltol, dt, maxiters, ckpt_every,
               tape_capacity, callback, on_replay=None) -> FitResult:
    """The fit with every iteration after the first replayed from one HIP graph.

    All state lives in device tensors the graph reads and writes in place:
    the parameters in units of ``scales``, Adam's moments and step count,
    each group's tape and gradient buffers, the loss and a status flag.  The
    first iteration runs eagerly: it builds the plans' device tables and
    workspaces, sizes each tape to 1.5 times its longest trajectory's
    attempts, and checks the statuses on the host.  Replays check nothing on
    the host; the largest status of any trajectory in any iteration is kept
    on the device and read once after the last replay (an error names it).
    Each iteration's loss is copied to a device array between replays.
    """
    import torch

    kw = dict(solver=solver, abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters, ckpt_every=ckpt_every,
              tape_capacity=tape_capacity)
    items = _engines(groups, parameters, **kw)
    scale = torch.as_tensor(np.asarray(scales, dtype=np.float64), device="cuda")
    if scale.numel() != len(parameters):
        raise ValueError("one scale per fitted parameter")
    us = [(init / scale).contiguous() for _, _, _, init in items]
    ms = [torch.zeros_like(u) for u in us]
    vs = [torch.zeros_like(u) for u in us]
    step = torch.zeros((), dtype=torch.float64, device="cuda")
    loss_now = torch.zeros((), dtype=torch.float64, device="cuda")
    worst = torch.zeros((), dtype=torch.int64, device="cuda")
    losses = torch.zeros(iterations, dtype=torch.float64, device="cuda")
    b1, b2, eps = _ADAM

    # first iteration, eager: tapes sized from the attempts it takes
    bufs = []
    for (eng, y0, tgt, _), u in zip(items, us):
        while True:
            buf = eng.buffers(int(tgt.numel()), y0)
            _, st, _ = eng.launch(u * scale, buf)
            codes = st.cpu().numpy()
            if np.any(codes == 7):
                eng.capacity = int(int(buf["tape"].n.max().item()) * 1.5) + 16
                continue
            if np.any(codes != 0):
                raise ValueError(f"group {len(bufs)}: trajectory {int(np.flatnonzero(codes)[0])} failed "
                                 f"(status {int(codes[codes != 0][0])})")
            need = int(int(buf["tape"].n.max().item()) * 1.5) + 16
            if need > eng.capacity:
                eng.capacity = need
                continue
            bufs.append(buf)
            break

    def iteration():
        step.add_(1.0)
        total = torch.zeros((), dtype=torch.float64, device="cuda")
        c1 = 1.0 - torch.pow(torch.full_like(step, b1), step)
        c2 = 1.0 - torch.pow(torch.full_like(step, b2), step)
        for (eng, _, tgt, _), buf, u, m, v in zip(items, bufs, us, ms, vs):
            counts, st, grad = eng.launch(u * scale, buf)
            torch.maximum(worst, st.abs().max().to(torch.int64), out=worst)
            resid = counts - tgt
            total = total + (resid * resid).sum()
            g = 2.0 * resid[:, None] * grad * scale
            m.mul_(b1).add_(g, alpha=1.0 - b1)
            v.mul_(b2).addcmul_(g, g, value=1.0 - b2)
            denom = (v.sqrt() / c2.sqrt()).add_(eps)
            u.addcdiv_(m, denom * (c1 / lr), value=-1.0)
        loss_now.copy_(total)

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    iteration()  # iteration 0, eager
    losses[0].copy_(loss_now)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            iteration()
    torch.cuda.current_stream().wait_stream(side)
    for it in range(1, iterations):
        graph.replay()
        losses[it].copy_(loss_now)
        if on_replay is not None:
            on_replay(it)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t_start
    code = int(worst.item())
    if code != 0:
        raise ValueError(f"a trajectory failed during the replayed iterations (status {code}; 7: the tape was "
                         "too short, raise tape_capacity)")
    loss_list = [float(x) for x in losses.cpu().numpy()]
    if callback is not None:
        for it, value in enumerate(loss_list):
            callback(it, value)
    params = [(u * scale).detach() for u in us]
    counts = [eng.forward(p, y0)[0].cpu().numpy() for (eng, y0, _, _), p in zip(items, params)]
    n_traj = sum(len(g.targets) for g in groups)
    tape_bytes = sum(buf["tape"].nbytes for buf in bufs)
    return FitResult([p.cpu().numpy() for p in params], loss_list, counts, seconds,
                     n_traj * iterations / seconds, tape_bytes)
