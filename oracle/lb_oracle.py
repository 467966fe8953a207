"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / reported CPU baseline.
The product path never imports it.

``Oracle(prepared)`` wraps ``liblb_oracle.so`` (built by ``oracle/Makefile``)
for one prepared problem.  Grid scans are expanded here into per-trajectory
base-slot overrides in the reference's row-major order, last axis fastest
(``rust/src/lindblad/ode_batch.rs:353-391``).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from centrex_tlf_amd.lindblad._abi import make_plan_desc, ptr

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblb_oracle.so"

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class LboSolveArgs(C.Structure):
    _fields_ = [
        ("solver", C.c_int32),
        ("t0", C.c_double),
        ("t1", C.c_double),
        ("abstol", C.c_double),
        ("reltol", C.c_double),
        ("dt", C.c_double),
        ("maxiters", C.c_int64),
        ("save_start", C.c_int32),
        ("saveat", _f64p),
        ("n_saveat", C.c_int64),
        ("output_kind", C.c_int32),
        ("store_trace", C.c_int32),
        ("n_selected", C.c_int32),
        ("selected_rows", _i32p),
        ("selected_cols", _i32p),
        ("n_weights", C.c_int32),
        ("weight_indices", _i32p),
        ("weight_values", _f64p),
        ("event_kind", C.c_int32),
        ("n_event_indices", C.c_int32),
        ("event_indices", _i32p),
        ("event_threshold", C.c_double),
        ("n_event_instr", C.c_int32),
        ("event_op", _i32p),
        ("event_index", _i32p),
        ("event_argc", _i32p),
        ("event_function", _i32p),
        ("event_re", _f64p),
        ("event_im", _f64p),
        ("event_t", _f64p),
        ("integral_out", _f64p),
        ("integral_save_start", C.c_int32),
        ("saveat_is_t1", C.c_int32),
    ]


OUT_KIND = {"populations": 0, "selected": 1, "integral": 2, "full": 3, "rate": 4}
INTEGRAL_OUTPUTS = {"weighted_integral", "photon_integral", "excited_population"}
RATE_OUTPUTS = {"weighted_rate", "photon_rate", "excited_population_rate"}


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(LIB_PATH))
        L.lbo_plan_create.restype = C.c_int
        L.lbo_plan_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.lbo_plan_destroy.argtypes = [C.c_void_p]
        L.lbo_eff_plan_create.restype = C.c_int
        L.lbo_eff_plan_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.lbo_last_error.restype = C.c_char_p
        L.lbo_plan_terms.restype = C.c_int64
        L.lbo_plan_terms.argtypes = [C.c_void_p]
        L.lbo_rhs.restype = C.c_int
        L.lbo_rhs.argtypes = [C.c_void_p, _f64p, C.c_double, C.c_int32, _i32p, _f64p, _f64p]
        L.lbo_hinit.restype = C.c_int
        L.lbo_hinit.argtypes = [C.c_void_p, _f64p, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int32, _i32p, _f64p, _f64p]
        L.lbo_eval_expression.restype = C.c_int
        L.lbo_eval_expression.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p,
                                          C.c_int32, _f64p, C.c_double, _f64p]
        L.lbo_output_width.restype = C.c_int64
        L.lbo_output_width.argtypes = [C.c_void_p, C.c_void_p]
        L.lbo_solve_batch.restype = C.c_int
        L.lbo_solve_batch.argtypes = [
            C.c_void_p, C.c_void_p, C.c_int64, _f64p, C.c_int32, _f64p, C.c_int32, _i32p, _f64p,
            C.c_int32, C.c_int64, _f64p, _f64p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _f64p,
        ]
        _lib = L
    return _lib


class OracleError(ValueError):
    pass


class Oracle:
    def __init__(self, prepared: Any):
        self.prepared = prepared
        self.n = prepared.layout.n
        self.dim = prepared.layout.packed_len
        desc, keep = make_plan_desc(prepared.flat_plan)
        handle = C.c_void_p()
        rc = lib().lbo_plan_create(C.byref(desc), C.byref(handle))
        del keep
        if rc != 0:
            raise OracleError(lib().lbo_last_error().decode())
        self._handle = handle

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.lbo_plan_destroy(h)
            self._handle = None

    def slot_indices(self, slots: Sequence[Any]) -> list[int]:
        names = self.prepared.slot_names
        out = []
        for s in slots:
            name = getattr(s, "name", str(s))
            out.append(names.index(name))
        return out

    def rhs(self, y: np.ndarray, t: float, overrides: dict[Any, complex] | None = None) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.dim, dtype=np.float64)
        slots, values = self._override_arrays(overrides)
        rc = lib().lbo_rhs(self._handle, ptr(y, C.c_double), float(t), len(slots),
                           ptr(slots, C.c_int32), ptr(values, C.c_double), ptr(out, C.c_double))
        if rc != 0:
            raise OracleError(lib().lbo_last_error().decode())
        return out

    def hinit(self, y: np.ndarray, t_span: tuple[float, float], abstol: float, reltol: float,
              overrides: dict[Any, complex] | None = None) -> float:
        """The starting step hinit chooses for y at t_span[0] (common.rs:72-118)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        slots, values = self._override_arrays(overrides)
        h = np.zeros(1, np.float64)
        rc = lib().lbo_hinit(self._handle, ptr(y, C.c_double), float(t_span[0]), float(t_span[1]),
                             float(abstol), float(reltol), len(slots), ptr(slots, C.c_int32),
                             ptr(values, C.c_double), ptr(h, C.c_double))
        if rc != 0:
            raise OracleError(lib().lbo_last_error().decode())
        return float(h[0])

    def _override_arrays(self, overrides):
        if not overrides:
            return np.zeros(0, np.int32), np.zeros(0, np.float64)
        slots = np.asarray(self.slot_indices(list(overrides)), dtype=np.int32)
        vals = np.asarray([complex(v) for v in overrides.values()], dtype=np.complex128)
        return slots, np.ascontiguousarray(vals.view(np.float64))

    def solve_batch(self, y0: np.ndarray, t_span: tuple[float, float], *, solver: str = "dopri5",
                    abstol: float = 1e-7, reltol: float = 1e-4, dt: float = 1e-8,
                    saveat: np.ndarray | None = None, save_start: bool = True,
                    maxiters: int = 100_000, output: str = "populations",
                    output_when: str = "final", output_indices=None, integral_weights=None,
                    parameter_slots: Sequence[Any] | None = None,
                    parameter_batch: np.ndarray | None = None,
                    t1_per_traj: np.ndarray | None = None, threads: int = 0,
                    stop_event: dict[str, Any] | None = None,
                    raise_on_error: bool = True, with_integral: bool = False) -> dict[str, Any]:
        """Reference-semantics batch solve; option translation of python_api.rs:589-617.

        ``stop_event`` is a spec of ``events.normalize_stop_event``; the root
        times come back as ``event_t`` (NaN where the event did not fire)."""
        y0 = np.asarray(y0, dtype=np.float64)
        shared = y0.ndim == 1
        n_traj = (parameter_batch.shape[0] if parameter_batch is not None else
                  (1 if shared else y0.shape[0]))
        if t1_per_traj is not None:
            n_traj = len(t1_per_traj)
        y0c = np.ascontiguousarray(y0.reshape(-1))
        final_integral = output_when == "final" and output in INTEGRAL_OUTPUTS
        sv = None if saveat is None else np.ascontiguousarray(saveat, dtype=np.float64)
        ss = bool(save_start)
        if output_when == "final" and not final_integral:
            sv = np.asarray([t_span[1]], dtype=np.float64)
        ss = ((output_when != "final") or final_integral) and ss
        kind = ("integral" if output in INTEGRAL_OUTPUTS else "rate" if output in RATE_OUTPUTS
                else output)
        a = LboSolveArgs()
        a.solver = {"dopri5": 0, "tsit5": 1, "fixed_dopri5": 2, "fixed_rk2": 3, "fixed_rk4": 4}[solver]
        a.t0, a.t1 = float(t_span[0]), float(t_span[1])
        a.abstol, a.reltol, a.dt = float(abstol), float(reltol), float(dt)
        a.maxiters = int(maxiters)
        a.save_start = int(ss)
        if t1_per_traj is not None and output_when == "final" and not final_integral:
            sv = None  # the plan {t1} of each trajectory
            a.saveat_is_t1 = 1
        a.saveat = ptr(sv, C.c_double)
        a.n_saveat = 0 if sv is None else sv.size
        a.output_kind = OUT_KIND[kind]
        a.store_trace = int(kind == "integral" and output_when == "saveat")
        keep = []
        if output == "selected":
            rows = np.asarray([i for i, _ in output_indices], dtype=np.int32)
            cols = np.asarray([j for _, j in output_indices], dtype=np.int32)
            keep += [rows, cols]
            a.n_selected = rows.size
            a.selected_rows, a.selected_cols = ptr(rows, C.c_int32), ptr(cols, C.c_int32)
        if integral_weights is not None:
            wi = np.asarray([int(i) for i, _ in integral_weights], dtype=np.int32)
            wv = np.asarray([float(w) for _, w in integral_weights], dtype=np.float64)
            keep += [wi, wv]
            a.n_weights = wi.size
            a.weight_indices, a.weight_values = ptr(wi, C.c_int32), ptr(wv, C.c_double)
        side = np.full(n_traj, np.nan)
        if with_integral:
            # the photon count of the reference's second solve (saveat None, save_start as given)
            a.integral_out = ptr(side, C.c_double)
            a.integral_save_start = int(bool(save_start))
        event_t = np.full(n_traj, np.nan)
        if stop_event is not None:
            a.event_t = ptr(event_t, C.c_double)
            if stop_event["kind"] == "population":
                idx = np.ascontiguousarray(stop_event["indices"], dtype=np.int32)
                keep.append(idx)
                a.event_kind, a.n_event_indices = 1, idx.size
                a.event_indices = ptr(idx, C.c_int32)
                a.event_threshold = float(stop_event["threshold"])
            else:
                ins = stop_event["expression"].instructions
                cols = [np.ascontiguousarray([q[k] for q in ins], dtype=np.int32) for k in range(4)]
                re = np.ascontiguousarray([q[4] for q in ins], dtype=np.float64)
                im = np.ascontiguousarray([q[5] for q in ins], dtype=np.float64)
                keep += cols + [re, im]
                a.event_kind, a.n_event_instr = 2, len(ins)
                a.event_op, a.event_index, a.event_argc, a.event_function = (ptr(c, C.c_int32) for c in cols)
                a.event_re, a.event_im = ptr(re, C.c_double), ptr(im, C.c_double)
        width = lib().lbo_output_width(self._handle, C.byref(a))
        if sv is not None:
            capacity = sv.size + 1
        elif a.saveat_is_t1:
            capacity = 2
        elif kind == "integral" and a.store_trace == 0:
            capacity = 1
        else:
            capacity = int(maxiters) + 2
        slots = np.zeros(0, np.int32)
        vals = np.zeros(0, np.float64)
        if parameter_slots is not None:
            slots = np.asarray(self.slot_indices(parameter_slots), dtype=np.int32)
            vals = np.ascontiguousarray(np.asarray(parameter_batch, dtype=np.complex128).view(np.float64))
        t1s = None if t1_per_traj is None else np.ascontiguousarray(t1_per_traj, dtype=np.float64)
        times = np.zeros(n_traj * capacity)
        values = np.zeros(n_traj * capacity * width)
        n_times = np.zeros(n_traj, dtype=np.int64)
        stats = np.zeros(n_traj * 4, dtype=np.int64)
        fail_x = np.zeros(n_traj)
        rc = lib().lbo_solve_batch(
            self._handle, C.byref(a), n_traj, ptr(y0c, C.c_double), int(shared), ptr(t1s, C.c_double),
            slots.size, ptr(slots, C.c_int32), ptr(vals, C.c_double), int(threads), capacity,
            ptr(times, C.c_double), ptr(values, C.c_double), ptr(n_times, C.c_int64),
            ptr(stats, C.c_int64), ptr(fail_x, C.c_double),
        )
        err = lib().lbo_last_error().decode() if rc != 0 else ""
        if rc != 0 and raise_on_error:
            raise OracleError(err)
        stats = stats.reshape(n_traj, 4)
        values = values.reshape(n_traj, capacity, width)
        times = times.reshape(n_traj, capacity)
        count = int(n_times.max()) if n_traj else 0
        result = {
            "times": times[:, :count],
            "values": values[:, :count, :],
            "n_times": n_times,
            "accepted": stats[:, 0],
            "rejected": stats[:, 1],
            "rhs_calls": stats[:, 2],
            "status": stats[:, 3],
            "fail_x": fail_x,
            "event_t": event_t,
            "integral": side if with_integral else None,
            "error": err,
            "width": width,
        }
        if output == "selected":
            result["values"] = result["values"].view(np.complex128)
        if output_when == "final" and count > 0:
            result["final"] = result["values"][:, -1, :]
        return result


class EffectiveOracle(Oracle):
    """The oracle on an effective-Hamiltonian plan (``lbo_eff_plan_create``):
    the same solvers, collectors and batch driver with the interpolated
    split-real RHS.  ``plan`` is a ``centrex_tlf_amd.lindblad.effective.EffectivePlan``
    (its host tables; no device needed)."""

    def __init__(self, plan: Any):
        class _Slots:
            slot_names = plan.slot_names

        self.prepared = _Slots()
        self.n = plan.n_states
        self.dim = plan.real_dim
        handle = C.c_void_p()
        rc = lib().lbo_eff_plan_create(C.byref(plan.desc), C.byref(handle))
        if rc != 0:
            raise OracleError(lib().lbo_last_error().decode())
        self._handle = handle


def eval_expression(instructions: Sequence[tuple], t: float = 0.0,
                    slots: Sequence[complex] = ()) -> complex:
    """Evaluate one program of (op, index, argc, function, value) records, the
    lb_plan_desc instruction format, with the oracle's expression machine."""
    n = len(instructions)
    op = np.asarray([q[0] for q in instructions], np.int32)
    idx = np.asarray([q[1] for q in instructions], np.int32)
    argc = np.asarray([q[2] for q in instructions], np.int32)
    fn = np.asarray([q[3] for q in instructions], np.int32)
    vals = np.asarray([complex(q[4]) for q in instructions], np.complex128)
    re = np.ascontiguousarray(vals.real)
    im = np.ascontiguousarray(vals.imag)
    sl = np.ascontiguousarray(np.asarray(list(slots) or [0j], np.complex128).view(np.float64))
    out = np.zeros(2, np.float64)
    rc = lib().lbo_eval_expression(n, ptr(op, C.c_int32), ptr(idx, C.c_int32), ptr(argc, C.c_int32),
                                   ptr(fn, C.c_int32), ptr(re, C.c_double), ptr(im, C.c_double),
                                   max(len(slots), 1), ptr(sl, C.c_double), float(t), ptr(out, C.c_double))
    if rc != 0:
        raise OracleError(lib().lbo_last_error().decode())
    return complex(out[0], out[1])


def grid_overrides(axes: Sequence[np.ndarray]) -> np.ndarray:
    """Cartesian product of axes, last axis fastest (ode_batch.rs:369-391)."""
    mesh = np.meshgrid(*[np.asarray(a, dtype=np.complex128) for a in axes], indexing="ij")
    return np.stack([m.reshape(-1) for m in mesh], axis=1)


def host_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
