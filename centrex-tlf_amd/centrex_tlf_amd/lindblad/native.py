"""ctypes binding of the HIP library (``liblindblad_gpu.so``, C ABI in ``include/centrex_lindblad.h``).

Device memory and streams come from PyTorch-ROCm; the solver itself is the
hand-written HIP code.  There is no CPU fallback: if the library or a HIP
device is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from ._abi import (LB_ADJOINT, LB_EVENT, LB_MODE_PERSISTENT, LB_MODE_PLAIN, LB_OUT, LB_SOLVER, LB_TAPE, LbSolveArgs,
                   make_plan_desc, ptr)

__all__ = ["DevicePlan", "NativeError", "Tape", "library", "library_path", "torch_device"]

LIB_NAME = "liblindblad_gpu.so"
_PKG_DIR = Path(__file__).resolve().parents[1]
_INFO_KEYS = [
    "n_states", "dim", "upper_len", "n_terms", "n_coefficients", "time_dependent", "n_published",
    "n_gather_terms", "n_self_terms", "slot_count", "gather_rows", "flops_per_rhs",
    "n_dynamic_coefficients", "lds_bytes", "kernel_variant", "threads_per_traj",
]


class NativeError(RuntimeError):
    pass


def library_path() -> Path:
    return _PKG_DIR / LIB_NAME


_lib: C.CDLL | None = None


def library() -> C.CDLL:
    """Load the HIP library (torch first, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (loads libamdhip64 before the library resolves it)

    path = library_path()
    if not path.exists():
        raise NativeError(
            f"{path} is not built; run __graft_entry__.build() (or make -C centrex-tlf_amd/csrc)"
        )
    L = C.CDLL(str(path))
    L.lb_plan_create.restype = C.c_int
    L.lb_plan_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.lb_plan_destroy.restype = None
    L.lb_plan_destroy.argtypes = [C.c_void_p]
    L.lb_last_error.restype = C.c_char_p
    L.lb_plan_specialize.restype = C.c_int
    L.lb_plan_specialize.argtypes = [C.c_void_p, C.c_int32]
    L.lb_plan_specialize_status.restype = C.c_int
    L.lb_plan_specialize_status.argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
    L.lb_plan_info.restype = C.c_int
    L.lb_plan_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.lb_plan_support_info.restype = C.c_int
    L.lb_plan_support_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]
    L.lb_rhs_eval.restype = C.c_int
    L.lb_rhs_eval.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                              C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p]
    L.lb_solve.restype = C.c_int
    L.lb_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lb_adjoint.restype = C.c_int
    L.lb_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lb_adjoint_noalloc.restype = C.c_int
    L.lb_adjoint_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lb_launch_info.restype = C.c_int
    L.lb_launch_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    L.lb_solve_group.restype = C.c_int
    L.lb_solve_group.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    L.lb_variant_info.restype = C.c_int
    L.lb_variant_info.argtypes = [C.c_int32] + [C.POINTER(C.c_int32)] * 3
    L.lb_eff_plan_create.restype = C.c_int
    L.lb_eff_plan_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.lb_eff_plan_destroy.restype = None
    L.lb_eff_plan_destroy.argtypes = [C.c_void_p]
    L.lb_eff_solve.restype = C.c_int
    L.lb_eff_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lb_eff_rhs_eval.restype = C.c_int
    L.lb_eff_rhs_eval.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]
    _lib = L
    return L


def kernel_variants() -> list[dict[str, int]]:
    """The compiled solver variants: slots per thread, largest workgroup, occupancy hint."""
    L = library()
    out = []
    i = 0
    while True:
        r, w, v = C.c_int32(), C.c_int32(), C.c_int32()
        if L.lb_variant_info(i, C.byref(r), C.byref(w), C.byref(v)) != 0:
            break
        out.append({"variant": i + 1, "slots_per_thread": r.value, "max_threads": w.value,
                    "waves_per_simd": v.value})
        i += 1
    return out


def _check(rc: int) -> None:
    if rc != 0:
        raise NativeError(library().lb_last_error().decode())


def torch_device():
    import torch

    if not torch.cuda.is_available():
        raise NativeError("no HIP device is available; the GPU solver needs an MI355X")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr() -> C.c_void_p:
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t) -> C.c_void_p | None:
    return None if t is None else C.c_void_p(t.data_ptr())


class GridAxes:
    """Scan axes uploaded once: complex values on the device plus their lengths."""

    def __init__(self, axes: Sequence[np.ndarray]):
        import torch

        arrays = [np.asarray(a, dtype=np.complex128).reshape(-1) for a in axes]
        if any(a.size == 0 for a in arrays):
            raise ValueError("scan axes must be non-empty")
        self.lengths = np.asarray([a.size for a in arrays], dtype=np.int64)
        self.count = int(np.prod(self.lengths))
        self.values = torch.view_as_real(
            torch.as_tensor(np.concatenate(arrays), device=torch_device())).contiguous()


class Tape:
    """Device arrays of an adjoint tape for B trajectories (see lb_solve_args):
    one record per step attempt (start ``t``, size ``h``, ``c`` = (error norm,
    proposed next step, fac_old), LB_TAPE_* flags ``f``), the state ``y`` at
    the start of every ``every``-th attempt, and ``n`` attempts per trajectory."""

    def __init__(self, n_traj: int, slots: int, cap: int, every: int = 16):
        import torch

        dev = torch_device()
        self.cap, self.every = int(cap), int(every)
        n_ckpt = (self.cap + self.every - 1) // self.every
        self.t = torch.zeros((n_traj, self.cap), dtype=torch.float64, device=dev)
        self.h = torch.zeros((n_traj, self.cap), dtype=torch.float64, device=dev)
        self.c = torch.zeros((n_traj, self.cap, 3), dtype=torch.float64, device=dev)
        self.f = torch.zeros((n_traj, self.cap), dtype=torch.int32, device=dev)
        self.y = torch.zeros((n_traj, n_ckpt, int(slots), 2), dtype=torch.float64, device=dev)
        self.n = torch.zeros(n_traj, dtype=torch.int32, device=dev)

    @property
    def nbytes(self) -> int:
        return sum(int(x.numel() * x.element_size()) for x in (self.t, self.h, self.c, self.f, self.y, self.n))

    def accepted(self) -> "Schedule":
        """The accepted step sizes in order (a replay schedule): the attempts
        flagged accepted, moved to the front of each row by a stable sort."""
        import torch

        live = torch.arange(self.cap, device=self.n.device)[None, :] < self.n[:, None].long()
        acc = live & ((self.f & LB_TAPE["accepted"]) != 0)
        order = torch.argsort((~acc).to(torch.int8), dim=1, stable=True)
        return Schedule(torch.gather(self.h, 1, order).contiguous(), acc.sum(dim=1).to(torch.int32), self.cap)


class Schedule:
    """Accepted step sizes (B, cap) and their counts (B,): a replay of a tape."""

    def __init__(self, h, n, cap: int):
        self.h, self.n, self.cap = h, n, int(cap)


class DevicePlan:
    """A lowered OBE system resident on the current HIP device."""

    def __init__(self, flat: dict[str, Any]):
        L = library()
        torch_device()
        desc, keep = make_plan_desc(flat)
        handle = C.c_void_p()
        rc = L.lb_plan_create(C.byref(desc), C.byref(handle))
        del keep
        _check(rc)
        self._handle = handle
        info = (C.c_int64 * 16)()
        _check(L.lb_plan_info(handle, info))
        self.info = dict(zip(_INFO_KEYS, [int(v) for v in info]))
        self.n = self.info["n_states"]
        self.dim = self.info["dim"]
        self.specialized = False
        self._spec_warned = False

    def specialize(self, enable: bool = True) -> "DevicePlan":
        """``lb_plan_specialize``: run the paired kernels of this plan with its
        time-dependent coefficient programs compiled in (bitwise the values of
        the expression machine).  Each (variant, solver) compiles on its first
        solve, about 2 s; a plan without time dependence is unchanged."""
        _check(library().lb_plan_specialize(self._handle, 1 if enable else 0))
        self.specialized = bool(enable)
        return self

    def specialize_failure(self) -> str | None:
        """``lb_plan_specialize_status``: the reason a specialised compilation
        of this plan failed (the precompiled kernels ran instead), or None."""
        buf = C.create_string_buffer(8192)
        rc = library().lb_plan_specialize_status(self._handle, buf, len(buf))
        if rc < 0:
            _check(rc)
        return buf.value.decode(errors="replace") if rc == 1 else None

    def support_of(self, y0) -> np.ndarray:
        """uint8 mask (dim) of the packed entries that are nonzero in any row of ``y0``."""
        a = np.asarray(y0.cpu() if hasattr(y0, "cpu") else y0)
        nz = (a != 0) if a.ndim == 1 else (a != 0).any(axis=0)
        return np.ascontiguousarray(nz.astype(np.uint8).reshape(-1))

    def support_info(self, support) -> dict[str, int]:
        """Plan facts of the restricted tables lb_solve uses for this initial support."""
        mask = np.ascontiguousarray(np.asarray(support, dtype=bool).astype(np.uint8).reshape(-1))
        if mask.size != self.dim:
            raise ValueError("support must have one entry per packed index")
        info = (C.c_int64 * 16)()
        _check(library().lb_plan_support_info(self._handle, ptr(mask, C.c_uint8), info))
        return dict(zip(_INFO_KEYS, [int(v) for v in info]))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.lb_plan_destroy(h)
            self._handle = None

    # -- batched RHS ---------------------------------------------------------
    def rhs(self, y, t, override_slots: Sequence[int] = (), override_values=None):
        """d(rho)/dt for a batch of packed states (torch float64 tensors on the device)."""
        import torch

        dev = torch_device()
        y = torch.as_tensor(y, dtype=torch.float64, device=dev).contiguous()
        if y.ndim == 1:
            y = y[None]
        B = y.shape[0]
        t = torch.as_tensor(t, dtype=torch.float64, device=dev).reshape(-1).expand(B).contiguous()
        slots = np.ascontiguousarray(override_slots, dtype=np.int32)
        ov = None
        if slots.size:
            ov = torch.view_as_real(torch.as_tensor(override_values, dtype=torch.complex128,
                                                    device=dev).reshape(B, slots.size)).contiguous()
        out = torch.zeros_like(y)
        _check(library().lb_rhs_eval(self._handle, B, _dptr(y), _dptr(t), slots.size,
                                     ptr(slots if slots.size else np.zeros(1, np.int32), C.c_int32),
                                     _dptr(ov), _dptr(out), _stream_ptr()))
        return out

    # -- integration ---------------------------------------------------------
    def launch_info(self, support=None, variant: int = 0, threads_per_traj: int = 0) -> dict[str, int]:
        """The persistent launch a solve would make: variant (1-based), threads,
        resident workgroups per CU and CUs (``lb_launch_info``)."""
        out = (C.c_int32 * 4)()
        sup = None
        if support is not None:
            sup = np.ascontiguousarray(np.asarray(support, dtype=bool).astype(np.uint8).reshape(-1))
            if sup.size != self.dim:
                raise ValueError("support must have one entry per packed index")
        _check(library().lb_launch_info(self._handle, None if sup is None else ptr(sup, C.c_uint8),
                                        int(variant), int(threads_per_traj), out))
        return {"variant": out[0], "threads_per_traj": out[1], "per_cu": out[2], "cus": out[3]}

    def solve(self, y0, t0: float, t1: float | Sequence[float], *, solver: str = "dopri5",
              abstol: float = 1e-7, reltol: float = 1e-4, dt: float = 1e-8, maxiters: int = 100_000,
              save_start: bool = True, output_kind: str = "populations",
              selected: Sequence[tuple[int, int]] | None = None,
              weights: Sequence[tuple[int, float]] | None = None,
              override_slots: Sequence[int] = (), override_values=None,
              grid_axes: Sequence[np.ndarray] | None = None, n_traj: int | None = None,
              mode: str = "persistent", threads_per_traj: int = 0, workgroups: int = 0,
              stream=None, with_integral: bool = False, variant: int = 0,
              restrict_support: bool = True, support=None, events=None,
              save_plan=None, trace: bool = False, capacity: int = 0, with_times: bool = False,
              tape=None, adjoint_grad=None, replay=None, queue_order=None, stop_event=None,
              adjoint_mode: str = "adaptive", noalloc: bool = False, _launch: bool = True):
        """Integrate a batch; returns (values, stats, fail_t) torch tensors on the device.

        ``y0`` is (dim,) shared or (B, dim).  Overrides are either per
        trajectory (``override_values`` of shape (B, n_override)) or a grid
        (``grid_axes``, one array per overridden slot, last axis fastest).
        With ``with_integral`` the weighted trapezoid is returned as a fourth
        tensor alongside a non-integral output, from the same pass.
        With ``restrict_support`` the solve runs only the density-matrix
        elements reachable from the nonzero entries of ``y0`` (the rest stay
        exactly zero); otherwise every element is integrated.

        Points are pushed at t0 (``save_start``) and every accepted step, or
        at the times of ``save_plan`` (dense output inside steps).  With
        ``trace`` every pushed point writes a row: the result is then
        ``(values[B, capacity, width], stats, fail_t, integral, counts, times)``
        (``times`` only with ``with_times``, ``integral`` only with
        ``with_integral``).

        ``tape`` (a :class:`Tape`) records every step attempt for the adjoint;
        with ``adjoint_grad`` (a (B, P) float64 tensor) the call runs the
        adjoint of that recorded tape instead of a solve (``lb_adjoint``):
        ``adjoint_mode`` "adaptive" differentiates the adaptive solve (the
        plan needs the time column, ``differentiate_flat_plan(...,
        with_time=True)``), "frozen" holds the accepted step sizes fixed.
        ``replay`` (a :class:`Tape` or :class:`Schedule`) integrates with the
        accepted step sizes of a recorded tape instead of the controller.
        ``noalloc`` runs the adjoint through ``lb_adjoint_noalloc`` (its
        workspace in the library's fixed device array; no allocation).

        ``stop_event`` (a spec from ``events.normalize_stop_event``) ends each
        trajectory at the first root of its event function; the root times
        (NaN where it did not fire) are appended to the result tuple.
        """
        import torch

        dev = torch_device()
        y0 = torch.as_tensor(y0, dtype=torch.float64, device=dev).contiguous()
        shared = y0.ndim == 1
        if support is not None:
            support = np.ascontiguousarray(np.asarray(support, dtype=bool).astype(np.uint8).reshape(-1))
            if support.size != self.dim:
                raise ValueError("support must have one entry per packed index")
        elif restrict_support:
            nz = (y0 != 0) if shared else (y0 != 0).any(dim=0)
            support = np.ascontiguousarray(nz.to("cpu").numpy().astype(np.uint8))
        slots = np.ascontiguousarray(override_slots, dtype=np.int32)
        axis_lengths = None
        axis_values = ov = None
        if isinstance(grid_axes, GridAxes):
            axis_lengths, axis_values = grid_axes.lengths, grid_axes.values
            B = grid_axes.count
        elif grid_axes is not None:
            g = GridAxes(grid_axes)
            axis_lengths, axis_values, B = g.lengths, g.values, g.count
        elif slots.size:
            ovt = torch.as_tensor(override_values, dtype=torch.complex128, device=dev)
            B = ovt.shape[0]
            ov = torch.view_as_real(ovt.reshape(B, slots.size)).contiguous()
        else:
            B = int(n_traj) if n_traj is not None else (1 if shared else y0.shape[0])
        if not shared and y0.shape[0] != B:
            raise ValueError(f"y0 has {y0.shape[0]} rows for {B} trajectories")
        t1_arr = None
        if torch.is_tensor(t1) and t1.ndim > 0:  # device tensors pass through (graph capture)
            t1_arr = t1.to(device=dev, dtype=torch.float64).contiguous()
            if t1_arr.numel() != B:
                raise ValueError("per-trajectory t1 must have one entry per trajectory")
            t1_scalar = float("nan")
        elif np.ndim(t1) > 0:
            t1_arr = torch.as_tensor(np.asarray(t1, dtype=np.float64), device=dev).contiguous()
            if t1_arr.numel() != B:
                raise ValueError("per-trajectory t1 must have one entry per trajectory")
            t1_scalar = float("nan")
        else:
            t1_scalar = float(t1)
        kind = LB_OUT[output_kind]
        if kind == LB_OUT["populations"]:
            width = self.n
        elif kind == LB_OUT["full"]:
            width = self.dim
        elif kind == LB_OUT["selected"]:
            width = 2 * len(selected)
        elif kind == LB_OUT["rate"] and not trace:
            raise ValueError("rate output needs trace=True")
        else:
            width = 1
        if trace:
            if capacity <= 0:
                raise ValueError("trace output needs a positive capacity")
            values = torch.zeros((B, int(capacity), width), dtype=torch.float64, device=dev)
        else:
            values = torch.full((B, width), float("nan"), dtype=torch.float64, device=dev)
        stats = torch.zeros((B, 4), dtype=torch.int64, device=dev)
        fail_t = torch.zeros(B, dtype=torch.float64, device=dev)
        keep: list[np.ndarray] = [slots]
        a = LbSolveArgs()
        a.n_traj = B
        a.solver = LB_SOLVER[solver]
        a.t0 = float(t0)
        a.t1 = t1_scalar
        a.t1_per_traj = _dptr(t1_arr)
        a.abstol, a.reltol, a.dt = float(abstol), float(reltol), float(dt)
        a.maxiters = int(maxiters)
        a.save_start = int(bool(save_start))
        a.y0 = _dptr(y0)
        a.y0_shared = int(shared)
        a.n_override = int(slots.size)
        a.override_slots = ptr(slots if slots.size else np.zeros(1, np.int32), C.c_int32)
        a.override_values = _dptr(ov)
        if axis_lengths is not None:
            keep.append(axis_lengths)
            a.n_axes = int(axis_lengths.size)
            a.axis_lengths = ptr(axis_lengths, C.c_int64)
            a.axis_values = _dptr(axis_values)
        a.output_kind = kind
        if selected is not None:
            rows = np.asarray([i for i, _ in selected], dtype=np.int32)
            cols = np.asarray([j for _, j in selected], dtype=np.int32)
            keep += [rows, cols]
            a.n_selected = rows.size
            a.selected_rows, a.selected_cols = ptr(rows, C.c_int32), ptr(cols, C.c_int32)
        if weights is not None:
            wi = np.asarray([int(i) for i, _ in weights], dtype=np.int32)
            wv = np.asarray([float(w) for _, w in weights], dtype=np.float64)
            keep += [wi, wv]
            a.n_weights = wi.size
            a.weight_indices, a.weight_values = ptr(wi, C.c_int32), ptr(wv, C.c_double)
        a.out_values = _dptr(values)
        a.out_stats = _dptr(stats)
        a.out_fail_t = _dptr(fail_t)
        integral = None
        if with_integral:
            integral = torch.full((B,), float("nan"), dtype=torch.float64, device=dev)
            a.out_integral = _dptr(integral)
        a.mode = LB_MODE_PLAIN if mode == "plain" else LB_MODE_PERSISTENT
        a.threads_per_traj = int(threads_per_traj)
        a.workgroups = int(workgroups)
        a.variant = int(variant)
        counts = times = plan_t = None
        if save_plan is not None:
            plan_t = torch.as_tensor(np.asarray(save_plan, dtype=np.float64).reshape(-1), device=dev).contiguous()
            if plan_t.numel() > 0:
                a.n_saveat = int(plan_t.numel())
                a.saveat = _dptr(plan_t)
        if trace:
            counts = torch.zeros(B, dtype=torch.int32, device=dev)
            a.trace = 1
            a.capacity = int(capacity)
            a.out_counts = _dptr(counts)
            if with_times:
                times = torch.full((B, int(capacity)), float("nan"), dtype=torch.float64, device=dev)
                a.out_times = _dptr(times)
        if support is not None:
            keep.append(support)
            a.y0_support = ptr(support, C.c_uint8)
        if replay is not None:  # integrate with the accepted step sizes of a recorded tape
            if isinstance(replay, Tape):
                replay = replay.accepted()
            keep.append(replay)
            a.replay_h, a.replay_n = _dptr(replay.h), _dptr(replay.n)
            a.tape_cap = int(replay.cap)
        if tape is not None:
            a.tape_t, a.tape_h = _dptr(tape.t), _dptr(tape.h)
            a.tape_c, a.tape_f = _dptr(tape.c), _dptr(tape.f)
            a.tape_y, a.tape_n = _dptr(tape.y), _dptr(tape.n)
            a.tape_cap, a.ckpt_every = int(tape.cap), int(tape.every)
        a.adjoint_mode = LB_ADJOINT[adjoint_mode]
        order_t = None
        if queue_order is not None:  # queue position -> trajectory (e.g. longest first)
            order_t = (queue_order.to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
                       if torch.is_tensor(queue_order) else
                       torch.as_tensor(np.asarray(queue_order, dtype=np.int32).reshape(-1), device=dev).contiguous())
            if order_t.numel() != B:
                raise ValueError("queue_order must have one entry per trajectory")
            a.queue_order = _dptr(order_t)
        if trace:
            result = (values, stats, fail_t, integral, counts, times)
        elif with_integral:
            result = (values, stats, fail_t, integral)
        else:
            result = (values, stats, fail_t)
        if stop_event is not None:
            event_t = torch.full((B,), float("nan"), dtype=torch.float64, device=dev)
            a.out_event_t = _dptr(event_t)
            if stop_event["kind"] == "population":
                idx = np.ascontiguousarray(stop_event["indices"], dtype=np.int32)
                keep.append(idx)
                a.event_kind = LB_EVENT["population"]
                a.n_event_indices = idx.size
                a.event_indices = ptr(idx, C.c_int32)
                a.event_threshold = float(stop_event["threshold"])
            else:
                ins = stop_event["expression"].instructions
                cols = [np.ascontiguousarray([q[k] for q in ins], dtype=np.int32) for k in range(4)]
                re = np.ascontiguousarray([q[4] for q in ins], dtype=np.float64)
                im = np.ascontiguousarray([q[5] for q in ins], dtype=np.float64)
                keep += cols + [re, im]
                a.event_kind = LB_EVENT["runtime_expression"]
                a.n_event_instr = len(ins)
                a.event_op, a.event_index, a.event_argc, a.event_function = (ptr(c, C.c_int32) for c in cols)
                a.event_re, a.event_im = ptr(re, C.c_double), ptr(im, C.c_double)
            result = result + (event_t,)
        if not _launch:
            if adjoint_grad is not None or events is not None or stop_event is not None:
                raise ValueError("grouped solves take no adjoint, timing events or stop event")
            return _Deferred(a, keep + [y0, t1_arr, ov, plan_t, order_t], result)
        sp = _stream_ptr() if stream is None else C.c_void_p(stream.cuda_stream)
        if events is not None:  # (start, end) torch events recorded right around the launch
            events[0].record(stream)
        if adjoint_grad is not None:
            a.out_grad = _dptr(adjoint_grad)
            entry = library().lb_adjoint_noalloc if noalloc else library().lb_adjoint
            _check(entry(self._handle, C.byref(a), sp))
            if events is not None:
                events[1].record(stream)
            return adjoint_grad
        _check(library().lb_solve(self._handle, C.byref(a), sp))
        if events is not None:
            events[1].record(stream)
        if self.specialized and not self._spec_warned:
            why = self.specialize_failure()
            if why is not None:
                import warnings

                self._spec_warned = True
                warnings.warn(f"specialised kernel not compiled, precompiled kernel used: {why}", stacklevel=2)
        return result


class _Deferred:
    """Arguments of a solve built but not launched (for ``solve_group``)."""

    def __init__(self, args, keep, result):
        self.args, self.keep, self.result = args, keep, result


def solve_group(calls: Sequence[tuple["DevicePlan", dict[str, Any]]], stream=None, events=None) -> list[tuple]:
    """Solve several plans concurrently in one call (``lb_solve_group``).

    ``calls`` is a list of ``(plan, kwargs)`` with the keyword arguments of
    :meth:`DevicePlan.solve` (``y0``, ``t0``, ``t1`` included); each plan may
    appear once.  Returns each call's result tuple, in order.  ``workgroups``
    in a call's kwargs bounds that group's share of the device.
    """
    if not calls:
        return []
    deferred = [plan.solve(_launch=False, **kw) for plan, kw in calls]
    n = len(calls)
    handles = (C.c_void_p * n)(*[plan._handle for plan, _ in calls])
    args = (LbSolveArgs * n)(*[d.args for d in deferred])
    sp = _stream_ptr() if stream is None else C.c_void_p(stream.cuda_stream)
    if events is not None:
        events[0].record(stream)
    _check(library().lb_solve_group(handles, args, n, sp))
    if events is not None:
        events[1].record(stream)
    return [d.result for d in deferred]


class EffectiveDevicePlan:
    """An effective-Hamiltonian plan (``lb_eff_plan``) on the current HIP device."""

    def __init__(self, desc, keep):
        L = library()
        torch_device()
        handle = C.c_void_p()
        _check(L.lb_eff_plan_create(C.byref(desc), C.byref(handle)))
        self._handle = handle
        self.real_dim = int(desc.real_dim)
        self.n_states = int(desc.n_states)
        del keep

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.lb_eff_plan_destroy(h)
            self._handle = None

    def rhs(self, y, t):
        """Batched RHS (``lb_eff_rhs_eval``): y (B, real_dim), t (B,); returns (dy, status)."""
        import torch

        dev = torch_device()
        y = torch.as_tensor(y, dtype=torch.float64, device=dev).reshape(-1, self.real_dim).contiguous()
        t = torch.as_tensor(t, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        out = torch.empty_like(y)
        status = torch.zeros(y.shape[0], dtype=torch.int32, device=dev)
        _check(library().lb_eff_rhs_eval(self._handle, y.shape[0], _dptr(y), _dptr(t), _dptr(out), _dptr(status),
                                         _stream_ptr()))
        return out, status

    def solve(self, y0, t0: float, t1: float, *, solver: str = "dopri5", abstol: float = 1e-9,
              reltol: float = 1e-7, dt: float = 1e-10, maxiters: int = 100_000, save_start: bool = True,
              save_plan=None, output_kind: str = "full", selected=None, weights=None,
              override_slots: Sequence[int] = (), override_values=None, n_traj: int = 1,
              capacity: int = 1, workgroups: int = 0):
        """``lb_eff_solve``: returns (values (B, capacity, width), stats, fail_t, counts, times) on the device."""
        import torch

        from ._abi import LB_EFF_OUT, LbEffSolveArgs

        dev = torch_device()
        y0 = torch.as_tensor(y0, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        if y0.numel() != self.real_dim:
            raise ValueError(f"expected state length {self.real_dim}, got {y0.numel()}")
        B = int(n_traj)
        kind = LB_EFF_OUT[output_kind]
        width = {0: self.n_states, 1: 2 * len(selected or ()), 2: 1, 3: self.real_dim}[kind]
        values = torch.full((B, int(capacity), width), float("nan"), dtype=torch.float64, device=dev)
        times = torch.full((B, int(capacity)), float("nan"), dtype=torch.float64, device=dev)
        stats = torch.zeros((B, 4), dtype=torch.int64, device=dev)
        fail_t = torch.zeros(B, dtype=torch.float64, device=dev)
        counts = torch.zeros(B, dtype=torch.int32, device=dev)
        keep = []
        a = LbEffSolveArgs()
        a.n_traj = B
        a.solver = LB_SOLVER[solver]
        a.t0, a.t1, a.abstol, a.reltol, a.dt = float(t0), float(t1), float(abstol), float(reltol), float(dt)
        a.maxiters = int(maxiters)
        a.save_start = int(bool(save_start))
        a.y0 = _dptr(y0)
        slots = np.ascontiguousarray(override_slots, dtype=np.int32)
        ov = None
        if slots.size:
            ov = torch.as_tensor(np.asarray(override_values, dtype=np.float64).reshape(B, slots.size),
                                 device=dev).contiguous()
            keep.append(slots)
            a.n_override = slots.size
            a.override_slots = ptr(slots, C.c_int32)
            a.override_values = _dptr(ov)
        a.output_kind = kind
        if selected is not None:
            rows = np.asarray([i for i, _ in selected], dtype=np.int32)
            cols = np.asarray([j for _, j in selected], dtype=np.int32)
            keep += [rows, cols]
            a.n_selected = rows.size
            a.selected_rows, a.selected_cols = ptr(rows, C.c_int32), ptr(cols, C.c_int32)
        if weights is not None:
            wi = np.asarray([int(i) for i, _ in weights], dtype=np.int32)
            wv = np.asarray([float(w) for _, w in weights], dtype=np.float64)
            keep += [wi, wv]
            a.n_weights = wi.size
            a.weight_indices, a.weight_values = ptr(wi, C.c_int32), ptr(wv, C.c_double)
        plan_t = None
        if save_plan is not None and np.size(save_plan):
            plan_t = torch.as_tensor(np.asarray(save_plan, dtype=np.float64).reshape(-1), device=dev).contiguous()
            a.n_saveat = int(plan_t.numel())
            a.saveat = _dptr(plan_t)
        a.capacity = int(capacity)
        a.out_values, a.out_times, a.out_counts = _dptr(values), _dptr(times), _dptr(counts)
        a.out_stats, a.out_fail_t = _dptr(stats), _dptr(fail_t)
        a.workgroups = int(workgroups)
        _check(library().lb_eff_solve(self._handle, C.byref(a), _stream_ptr()))
        return values, stats, fail_t, counts, times
