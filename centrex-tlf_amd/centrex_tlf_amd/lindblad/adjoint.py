"""Gradients of photon counts through the discrete adjoint on the GPU.

``DifferentiablePhotonCount`` wraps a differentiable device plan: the
prepared problem's plan plus the parameter- and time-derivative programs of
its coefficients (``derivatives.differentiate_flat_plan``) and the
transposed term tables.  Its forward pass is one batched solve that records
a tape (every step attempt, with the controller's inputs and decision, and a
checkpoint every ``ckpt_every`` attempts); its backward pass is one launch of
the adjoint kernel (``lb_adjoint``), which recomputes each checkpoint segment
and sweeps the Dopri5/Tsit5 stage adjoints backwards.

With ``mode="adaptive"`` (the default) the gradient is that of the adaptive
solve itself: the error norm, the step-size controller, rejected attempts,
the last-step clip and the trapezoid's step lengths are differentiated, so it
is the derivative the finite differences of the forward see.  ``mode="frozen"``
holds the accepted step sizes fixed (the derivative of ``replay``).
``photon_count`` exposes it as a ``torch.autograd.Function``.

The reference has no gradients; the objective is its photon-count output
(``output="photon_integral"``: trapezoid of sum_k w_k rho_kk over t0 and every
accepted step, ``rust/src/ode/output.rs:253-335``).
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import numpy as np

from . import _driver as D
from .derivatives import differentiate_flat_plan
from .native import DevicePlan, Tape

__all__ = ["DifferentiablePhotonCount", "photon_count"]


class DifferentiablePhotonCount:
    """Photon count I_b(p_b) of a batch and its parameter gradient.

    ``parameters`` names real base parameters (e.g. ``["delta0", "Omega0",
    "theta", "sigma"]``); each trajectory b has its own values ``p[b]``.
    ``t_span`` is (t0, t1) or (t0, per-trajectory end times); ``fixed`` maps
    further base parameters to per-trajectory values that are overridden but
    not differentiated (e.g. a beam width per trajectory).
    """

    def __init__(self, prepared: Any, parameters: Sequence[Any], integral_weights: Sequence[tuple[int, float]],
                 t_span: Sequence[float], *, solver: str = "dopri5", abstol: float = 1e-7,
                 reltol: float = 1e-4, dt: float = 1e-8, maxiters: int = 100_000, save_start: bool = True,
                 ckpt_every: int = 16, tape_capacity: int = 1024, fixed: dict[str, Any] | None = None,
                 mode: str = "adaptive"):
        if solver not in D.GPU_SOLVERS:
            raise NotImplementedError("the adjoint supports 'dopri5' and 'tsit5'")
        if mode not in ("adaptive", "frozen"):
            raise ValueError("mode must be 'adaptive' or 'frozen'")
        self.mode = mode
        self.prepared = prepared
        self.names = [D.parameter_slot_name(p) for p in parameters]
        self.slots = D.parameter_slot_indices(prepared, parameters)
        self.weights = [(int(i), float(w)) for i, w in integral_weights]
        if np.ndim(t_span[1]) > 0:
            t1 = np.ascontiguousarray(t_span[1], dtype=np.float64).reshape(-1)
            if not np.all(np.isfinite(t1)):
                raise ValueError("end times must be finite")
            self.t_span = (float(t_span[0]), t1)
        else:
            self.t_span = D.normalize_t_span(t_span)
        fixed = fixed or {}
        self.fixed_slots = D.parameter_slot_indices(prepared, list(fixed.keys())) if fixed else []
        self.fixed_values = (np.stack([np.asarray(v, dtype=np.complex128).reshape(-1) for v in fixed.values()], axis=1)
                             if fixed else None)
        self.kw = dict(solver=solver, abstol=float(abstol), reltol=float(reltol), dt=float(dt),
                       maxiters=int(maxiters), save_start=bool(save_start))
        self.every = int(ckpt_every)
        self.capacity = int(tape_capacity)
        self.plan = DevicePlan(differentiate_flat_plan(prepared.flat_plan, self.slots, with_time=True))

    def _common(self, params, y0):
        import torch

        p = torch.as_tensor(params, dtype=torch.float64, device="cuda")
        if p.ndim != 2 or p.shape[1] != len(self.slots):
            raise ValueError(f"params must have shape (B, {len(self.slots)})")
        y = torch.as_tensor(np.asarray(y0, dtype=np.float64) if not torch.is_tensor(y0) else y0,
                            dtype=torch.float64, device="cuda")
        support = self.plan.support_of(y)
        slots = self.plan.support_info(support)["slot_count"]
        return p, y, support, slots

    def _overrides(self, p):
        """Override slots and (B, P + F) complex values: the parameters, then the fixed ones."""
        import torch

        vals = p.to(torch.complex128)
        if self.fixed_values is None:
            return self.slots, vals
        fv = torch.as_tensor(self.fixed_values, device=vals.device)
        if fv.shape[0] != vals.shape[0]:
            raise ValueError("fixed parameter values need one row per trajectory")
        return self.slots + self.fixed_slots, torch.cat([vals, fv], dim=1)

    def forward(self, params, y0, events=None):
        """Photon counts (B,) and the tape of the solve (``events``: a (start, end)
        pair of torch events recorded around the solver launch)."""
        import torch

        p, y, support, slots = self._common(params, y0)
        cap = self.capacity
        while True:
            tape = Tape(p.shape[0], slots, cap, self.every)
            vals, stats, fail_t = self.plan.solve(
                y, *self.t_span, output_kind="integral", weights=self.weights, override_slots=self._overrides(p)[0],
                override_values=self._overrides(p)[1], support=support, tape=tape, events=events, **self.kw)
            st = stats[:, 3].cpu().numpy()
            if np.any(st == 7):
                # the tape was too short: grow it (with a margin the next calls keep)
                cap = int(int(tape.n.max().item()) * 1.25) + 16
                self.capacity = cap
                continue
            bad = np.flatnonzero(st != 0)
            if bad.size:
                i = int(bad[0])
                raise ValueError("trajectory {}: {}".format(i, D.trajectory_message(
                    int(st[i]), float(fail_t[i].item()), maxiters=self.kw["maxiters"], n_plan=0)))
            return vals[:, 0], tape

    def buffers(self, n_traj: int, y0) -> dict[str, Any]:
        """Device buffers for :meth:`launch`: the packed initial state, its
        support, a tape of the engine's capacity, the per-trajectory end times
        and fixed parameter values on the device, and the gradient output."""
        import torch

        y = torch.as_tensor(np.asarray(y0, dtype=np.float64) if not torch.is_tensor(y0) else y0,
                            dtype=torch.float64, device="cuda")
        support = self.plan.support_of(y)
        slots = self.plan.support_info(support)["slot_count"]
        t1 = self.t_span[1]
        if np.ndim(t1) > 0:
            if len(t1) != n_traj:
                raise ValueError("one end time per trajectory")
            t1 = torch.as_tensor(np.asarray(t1, dtype=np.float64), device="cuda")
        fixed = None
        if self.fixed_values is not None:
            if self.fixed_values.shape[0] != n_traj:
                raise ValueError("fixed parameter values need one row per trajectory")
            fixed = torch.as_tensor(self.fixed_values, device="cuda")
        return {"y": y, "support": support, "t1": t1, "fixed": fixed,
                "tape": Tape(n_traj, slots, self.capacity, self.every),
                "grad": torch.empty((n_traj, len(self.slots)), dtype=torch.float64, device="cuda")}

    def launch(self, p, buf: dict[str, Any]):
        """Photon counts (B,), per-trajectory status (B,) and dI/dp (B, P) of
        ``p`` (a (B, P) float64 device tensor): the forward solve recording
        ``buf["tape"]``, then the adjoint.  Nothing is read back to the host
        and nothing depends on a host value computed here, so a stream
        capture of this call replays it exactly; statuses (LB_TRAJ_*, 7 when
        the tape was too short) are for the caller to check."""
        import torch

        vals = p.to(torch.complex128)
        slots = self.slots
        if buf["fixed"] is not None:
            vals = torch.cat([vals, buf["fixed"]], dim=1)
            slots = self.slots + self.fixed_slots
        tape = buf["tape"]
        out, stats, _ = self.plan.solve(
            buf["y"], self.t_span[0], buf["t1"], output_kind="integral", weights=self.weights,
            override_slots=slots, override_values=vals, support=buf["support"], tape=tape, **self.kw)
        self.plan.solve(buf["y"], self.t_span[0], buf["t1"], output_kind="integral", weights=self.weights,
                        override_slots=slots, override_values=vals, support=buf["support"], tape=tape,
                        adjoint_grad=buf["grad"], adjoint_mode=self.mode, **self.kw)
        return out[:, 0], stats[:, 3], buf["grad"]

    def replay(self, params, y0, tape):
        """Photon counts (B,) integrated with the accepted step sizes recorded
        in ``tape`` (no step-size control): the map the frozen adjoint
        differentiates."""
        import torch

        p, y, support, _ = self._common(params, y0)
        vals, stats, _ = self.plan.solve(
            y, *self.t_span, output_kind="integral", weights=self.weights, override_slots=self._overrides(p)[0],
            override_values=self._overrides(p)[1], support=support, replay=tape, **self.kw)
        if int(stats[:, 3].abs().max().item()) != 0:
            raise ValueError("replay of the recorded step sizes failed")
        return vals[:, 0]

    def gradient(self, params, y0, tape, variant: int = 0, events=None, mode: str | None = None,
                 noalloc: bool = False):
        """dI_b/dp_b (B, P) for a tape recorded by ``forward`` with the same params.

        ``variant`` = k + 1 forces adjoint kernel variant k (0: the library's
        choice); ``mode`` overrides the engine's ("adaptive" or "frozen");
        ``noalloc`` uses ``lb_adjoint_noalloc`` (no device allocation).
        """
        import torch

        p, y, support, _ = self._common(params, y0)
        grad = torch.empty((p.shape[0], len(self.slots)), dtype=torch.float64, device="cuda")
        return self.plan.solve(
            y, *self.t_span, output_kind="integral", weights=self.weights, override_slots=self._overrides(p)[0],
            override_values=self._overrides(p)[1], support=support, tape=tape, adjoint_grad=grad,
            variant=int(variant), events=events, adjoint_mode=mode or self.mode, noalloc=noalloc, **self.kw)


def _autograd_function():
    import torch

    class PhotonCount(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params, engine, y0):
            counts, tape = engine.forward(params.detach(), y0)
            ctx.save_for_backward(params.detach())
            ctx.engine, ctx.tape, ctx.y0 = engine, tape, y0
            return counts

        @staticmethod
        def backward(ctx, grad_counts):
            (params,) = ctx.saved_tensors
            dIdp = ctx.engine.gradient(params, ctx.y0, ctx.tape)
            return grad_counts[:, None] * dIdp, None, None

    return PhotonCount


_FN = None


def photon_count(engine: DifferentiablePhotonCount, params, y0):
    """Differentiable photon counts: ``photon_count(engine, params (B, P), y0) -> (B,)``."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(params, engine, y0)
