"""Effective-Hamiltonian Lindblad path on the GPU.

Mirrors the reference's Rust bindings for this path
(``rust/src/effective_lindblad/python_api.rs``): ``prepare_effective_lindblad_plan``
takes the plan dictionary ``centrex_tlf/effective_hamiltonian/rust_plan.py:42-201``
builds (split-real interpolated operators, field grid, slots, parameter
graph) and uploads it; ``solve_effective_lindblad`` and
``solve_effective_lindblad_batch`` take the same arguments and return the
same ``(times, values, width[, time_count, stats])`` as the Rust functions;
``debug_effective_lindblad_rhs`` returns the same diagnostic dictionary.

The state is y = [Re rho.flat, Im rho.flat] (2 n^2 reals) and
dy = A_comb(x) y + (Omega/2) A_opt(x) y + delta A_det(x) y, with each operator
entry a cubic in x on each field-grid interval (``rhs.rs:86-138``).

``effective_operator_payload`` restates the operator-table builder of
``rust_plan.py:70-133`` (union sparsity over the grid, linear or pchip
slopes, interval-major coefficients) for building plans from dense split-real
superoperators; ``split_real_superoperator`` / ``split_real_rho`` are the
layout conversions of ``rust_plan.py:18-39``.
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Sequence
from typing import Any

import numpy as np

from . import _driver as D
from ._abi import LbEffPlanDesc, ptr
from .ir import lower_parameter_graph

__all__ = [
    "EffectivePlan",
    "debug_effective_lindblad_rhs",
    "effective_operator_payload",
    "prepare_effective_lindblad_plan",
    "solve_effective_lindblad",
    "solve_effective_lindblad_batch",
    "split_real_rho",
    "split_real_superoperator",
    "split_real_to_rho",
]

_INTEGRAL_OUTPUTS = {"weighted_integral", "photon_integral", "excited_population"}
_EFF_SOLVERS = {"dopri5", "tsit5"}


def split_real_superoperator(L: np.ndarray) -> np.ndarray:
    """Complex superoperator on rho.flat -> real operator on [Re, Im] (``rust_plan.py:18-27``)."""
    n2 = L.shape[0]
    A, B = np.real(L), np.imag(L)
    out = np.zeros((2 * n2, 2 * n2), dtype=np.float64)
    out[:n2, :n2] = A
    out[:n2, n2:] = -B
    out[n2:, :n2] = B
    out[n2:, n2:] = A
    return out


def split_real_rho(rho: np.ndarray) -> np.ndarray:
    flat = np.asarray(rho).reshape(-1)
    return np.concatenate([np.real(flat), np.imag(flat)]).astype(np.float64)


def split_real_to_rho(y: np.ndarray, n_states: int) -> np.ndarray:
    n2 = n_states * n_states
    return (np.asarray(y[:n2]) + 1j * np.asarray(y[n2:])).reshape(n_states, n_states)


def effective_operator_payload(matrices: Sequence[np.ndarray], grid: np.ndarray,
                               interpolation: str = "linear") -> dict[str, Any]:
    """CSR over the union sparsity of ``matrices`` (one real matrix per grid
    point) with per-interval cubic coefficients (c0, c1, c2, c3) in
    ``x - grid[k]``: linear slopes, or the reference's pchip slopes
    (harmonic-mean interior slopes, one-sided ends; ``rust_plan.py:70-133``)."""
    mats = [np.asarray(m, dtype=np.float64) for m in matrices]
    grid = np.asarray(grid, dtype=np.float64)
    dim = mats[0].shape[0]
    mask = np.zeros((dim, dim), dtype=bool)
    for m in mats:
        mask |= m != 0.0
    row_ptrs = [0]
    col_indices: list[int] = []
    for i in range(dim):
        cols = np.flatnonzero(mask[i])
        col_indices.extend(int(c) for c in cols)
        row_ptrs.append(len(col_indices))
    nnz = len(col_indices)
    rows_of = np.repeat(np.arange(dim), np.diff(row_ptrs))
    cols_arr = np.asarray(col_indices, dtype=np.int64)
    values = np.stack([m[rows_of, cols_arr] for m in mats], axis=1) if nnz else np.zeros((0, grid.size))
    h = np.diff(grid)
    n_g = grid.size
    coeffs = np.zeros((n_g - 1, nnz, 4), dtype=np.float64)
    # one entry per row of `values`; every operation elementwise, as per entry
    delta = np.diff(values, axis=1) / h
    coeffs[:, :, 0] = values[:, :-1].T
    if interpolation == "pchip":
        d = np.zeros((nnz, n_g))
        d[:, 0] = delta[:, 0]
        d[:, -1] = delta[:, -1]
        for k in range(1, n_g - 1):
            w1 = 2.0 * h[k] + h[k - 1]
            w2 = h[k] + 2.0 * h[k - 1]
            same = delta[:, k - 1] * delta[:, k] > 0.0
            with np.errstate(divide="ignore", invalid="ignore"):
                hm = (w1 + w2) / (w1 / delta[:, k - 1] + w2 / delta[:, k])
            d[:, k] = np.where(same, hm, 0.0)
        coeffs[:, :, 1] = d[:, :-1].T
        coeffs[:, :, 2] = ((3.0 * delta - 2.0 * d[:, :-1] - d[:, 1:]) / h).T
        coeffs[:, :, 3] = ((d[:, :-1] + d[:, 1:] - 2.0 * delta) / (h * h)).T
    else:
        coeffs[:, :, 1] = delta.T
    return {"row_ptrs": np.asarray(row_ptrs, dtype=np.int64), "col_indices": cols_arr,
            "pchip_coeffs": coeffs.reshape(-1), "grid": grid, "nnz": nnz}


def _instructions(expr: Any) -> list[tuple[int, int, int, int, float, float]]:
    """Rows of one compiled expression: this package's CompiledExpression or a
    reference payload (``{"instructions": [{"op": .., "index": .., ...}]}``)."""
    if hasattr(expr, "instructions") and not isinstance(expr, dict):
        return [tuple(r) for r in expr.instructions]
    rows = []
    for ins in expr["instructions"]:
        if isinstance(ins, dict):
            rows.append((int(ins.get("op", 0)), int(ins.get("index", 0)), int(ins.get("argc", 0)),
                         int(ins.get("function", 0)), float(ins.get("re", 0.0)), float(ins.get("im", 0.0))))
        else:
            rows.append(tuple(ins))
    return rows


def _graph_arrays(graph: Any) -> dict[str, Any]:
    """Slots, base values and compound bytecode from a LindbladParameters
    object, this package's lowered graph or a reference graph payload."""
    if hasattr(graph, "slot_index_by_name"):
        graph = lower_parameter_graph(graph)
    base = [complex(v) for v in graph["base_values"]]
    comps = []
    for c in graph["compounds"]:
        if isinstance(c, dict):
            comps.append((int(c["slot"]), _instructions(c["expression"])))
        else:
            slot, expr = c
            comps.append((int(slot), _instructions(expr)))
    rows = [r for _, ins in comps for r in ins]
    ptrs = np.zeros(len(comps) + 1, dtype=np.int32)
    np.cumsum([len(ins) for _, ins in comps], out=ptrs[1:])
    ins = np.asarray(rows, dtype=np.float64).reshape(-1, 6) if rows else np.zeros((0, 6))
    return {
        "slot_names": list(graph["slot_names"]),
        "base": np.asarray(base, dtype=np.complex128),
        "compound_slots": np.asarray([s for s, _ in comps], dtype=np.int32),
        "expr_ptrs": ptrs,
        "ins": ins,
    }


class EffectivePlan:
    """A prepared effective-Hamiltonian plan (``EffectiveLindbladPlan``) on the GPU."""

    def __init__(self, plan_dict: dict[str, Any], *, device: bool = True):
        from .native import EffectiveDevicePlan

        def need(key):
            if key not in plan_dict:
                raise ValueError(f"missing key: {key}")
            return plan_dict[key]

        self.n_states = int(need("n_states"))
        self.real_dim = int(need("real_dim"))
        self.n_grid = int(need("n_grid"))
        self.field_grid = np.ascontiguousarray(need("field_grid"), dtype=np.float64)
        if self.field_grid.size != self.n_grid:
            raise ValueError("field_grid length mismatch")
        self.operator_interpolation = plan_dict.get("operator_interpolation", "linear")
        self.field_slot = int(need("field_coordinate_slot"))
        self.rabi_slot = int(need("rabi_rate_slot"))
        self.detuning_slot = int(need("detuning_slot"))
        if "parameter_graph" not in plan_dict:
            raise ValueError("missing 'parameter_graph'")
        g = _graph_arrays(plan_dict["parameter_graph"])
        self.slot_names = g["slot_names"]
        keep: list[Any] = [self.field_grid]
        desc = LbEffPlanDesc()
        desc.n_states, desc.real_dim, desc.n_grid = self.n_states, self.real_dim, self.n_grid
        desc.field_grid = ptr(self.field_grid, C.c_double)
        self.operators = []
        for o, key in enumerate(("sparse_combined", "sparse_opt", "sparse_det")):
            if key not in plan_dict:
                raise ValueError(f"missing sparse operator: {key}")
            op = plan_dict[key]
            rows = np.ascontiguousarray(op["row_ptrs"], dtype=np.int64)
            cols = np.ascontiguousarray(op["col_indices"], dtype=np.int64)
            coef = np.ascontiguousarray(op["pchip_coeffs"], dtype=np.float64)
            nnz = int(op["nnz"])
            if rows.size != self.real_dim + 1 or cols.size != nnz or coef.size != (self.n_grid - 1) * nnz * 4:
                raise ValueError(f"{key}: table sizes do not match real_dim, nnz and n_grid")
            keep += [rows, cols, coef]
            self.operators.append({"row_ptrs": rows, "col_indices": cols, "pchip_coeffs": coef, "nnz": nnz})
            desc.row_ptrs[o] = rows.ctypes.data
            desc.col_indices[o] = cols.ctypes.data
            desc.coeffs[o] = coef.ctypes.data
            desc.nnz[o] = nnz
        desc.field_slot, desc.rabi_slot, desc.detuning_slot = self.field_slot, self.rabi_slot, self.detuning_slot
        base_re = np.ascontiguousarray(g["base"].real)
        base_im = np.ascontiguousarray(g["base"].imag)
        ins = g["ins"]
        cols6 = [np.ascontiguousarray(ins[:, k].astype(np.int32)) for k in range(4)]
        ins_re, ins_im = np.ascontiguousarray(ins[:, 4]), np.ascontiguousarray(ins[:, 5])
        keep += [base_re, base_im, g["compound_slots"], g["expr_ptrs"], ins_re, ins_im] + cols6
        desc.n_slots = len(self.slot_names)
        desc.n_base = g["base"].size
        desc.base_re, desc.base_im = ptr(base_re, C.c_double), ptr(base_im, C.c_double)
        desc.n_compounds = g["compound_slots"].size
        desc.compound_slots = ptr(g["compound_slots"], C.c_int32)
        desc.expr_ptrs = ptr(g["expr_ptrs"], C.c_int32)
        desc.instr_op, desc.instr_index, desc.instr_argc, desc.instr_function = (ptr(c, C.c_int32) for c in cols6)
        desc.instr_re, desc.instr_im = ptr(ins_re, C.c_double), ptr(ins_im, C.c_double)
        self.graph = g
        self.desc, self._keep = desc, keep  # the oracle builds its own plan from these tables
        self.device = EffectiveDevicePlan(desc, keep) if device else None


def prepare_effective_lindblad_plan(plan_dict: dict[str, Any]) -> EffectivePlan:
    """``prepare_effective_lindblad_plan_py`` (``python_api.rs:16-26``)."""
    return EffectivePlan(plan_dict)


def _kind(output: str) -> str:
    if output in _INTEGRAL_OUTPUTS:
        return "integral"
    if output in {"populations", "selected"}:
        return output
    return "full"


def _fail(stats: np.ndarray, fail_t: np.ndarray, plan: EffectivePlan, maxiters: int, n_plan: int,
          batch: bool) -> None:
    bad = np.flatnonzero(stats[:, 3] != 0)
    if not bad.size:
        return
    i = int(bad[0])
    status = int(stats[i, 3])
    if status in (10, 11):
        word, bound = ("below", "minimum") if status == 10 else ("above", "maximum")
        g = plan.field_grid[0] if status == 10 else plan.field_grid[-1]
        msg = f"field coordinate {_rust_e6(fail_t[i])} is {word} the operator grid {bound} {_rust_e6(g)}"
    else:
        msg = D.trajectory_message(status, float(fail_t[i]), maxiters=maxiters, n_plan=n_plan)
    raise ValueError(f"trajectory {i}: {msg}" if batch else msg)


def _rust_e6(x: float) -> str:
    """Rust's ``{:.6e}``: six fraction digits, unpadded exponent."""
    mant, exp = f"{float(x):.6e}".split("e")
    return f"{mant}e{int(exp)}"


def _run(plan: EffectivePlan, y0: np.ndarray, t0: float, t1: float, *, abstol: float, reltol: float, dt: float,
         saveat, save_start: bool, maxiters: int, solver: str, output: str, output_indices, integral_weights,
         slot_indices: Sequence[int], parameter_batch, trajectory_count: int, batch: bool):
    if solver not in _EFF_SOLVERS:
        raise NotImplementedError(f"solver={solver!r} is not available on the GPU effective path; use 'dopri5' "
                                  "or 'tsit5'")
    kind = _kind(output)
    if kind == "selected" and output_indices is None:
        raise ValueError("output='selected' requires output_indices")
    if kind == "integral" and integral_weights is None:
        raise ValueError(f"output='{output}' requires integral_weights")
    sv = None if saveat is None else np.asarray(saveat, dtype=np.float64).reshape(-1)
    save_plan = D.build_save_plan(sv, t0, t1, bool(save_start))
    if kind == "integral":
        capacity = 1
    elif save_plan is not None:
        capacity = max(1, save_plan.size)
    else:
        capacity = int(maxiters) + 2
    values, stats, fail_t, counts, times = plan.device.solve(
        y0, t0, t1, solver=solver, abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters,
        save_start=save_start, save_plan=save_plan, output_kind=kind,
        selected=None if output_indices is None else [(int(i), int(j)) for i, j in output_indices],
        weights=integral_weights, override_slots=slot_indices, override_values=parameter_batch,
        n_traj=trajectory_count, capacity=capacity)
    stats, fail_t = stats.cpu().numpy(), fail_t.cpu().numpy()
    _fail(stats, fail_t, plan, maxiters, 0 if save_plan is None else save_plan.size, batch)
    return values.cpu().numpy(), stats, counts.cpu().numpy(), times.cpu().numpy(), kind


def solve_effective_lindblad(plan: EffectivePlan, y0, t0: float, t1: float, abstol: float, reltol: float,
                             dt: float, saveat=None, save_start: bool = True, maxiters: int = 100_000,
                             solver: str = "dopri5", output: str = "full", output_indices=None,
                             integral_weights=None):
    """``solve_effective_lindblad_py`` (``python_api.rs:28-91``): (times, values, width)."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1)
    values, stats, counts, times, kind = _run(
        plan, y0, float(t0), float(t1), abstol=abstol, reltol=reltol, dt=dt, saveat=saveat,
        save_start=save_start, maxiters=maxiters, solver=solver, output=output, output_indices=output_indices,
        integral_weights=integral_weights, slot_indices=(), parameter_batch=None, trajectory_count=1,
        batch=False)
    count = int(counts[0])
    width = values.shape[2]
    t = times[0, :count]
    v = values[0, :count]
    if kind == "selected":
        return t, (v[..., 0::2] + 1j * v[..., 1::2]).reshape(-1), width // 2
    return t, v.reshape(-1), width


def solve_effective_lindblad_batch(plan: EffectivePlan, packed_rho0, t0: float, t1: float, abstol: float,
                                   reltol: float, dt: float, saveat=None, save_start: bool = True,
                                   maxiters: int = 100_000, solver: str = "dopri5", output: str = "populations",
                                   output_indices=None, integral_weights=None, parameter_slot_indices=None,
                                   parameter_batch=None, trajectory_count: int = 1, parallel: bool = True,
                                   threads: int | None = None):
    """``solve_effective_lindblad_batch_py`` (``python_api.rs:152-250``, ``solver.rs:142-281``):
    (times, flat values, width, time_count, stats).  ``parallel``/``threads`` have no effect on the GPU."""
    y0 = np.ascontiguousarray(packed_rho0, dtype=np.float64).reshape(-1)
    slots = list(parameter_slot_indices or [])
    pb = None
    if parameter_batch is not None:
        pb = np.ascontiguousarray(parameter_batch, dtype=np.float64)
        if pb.shape != (trajectory_count, len(slots)):
            raise ValueError(f"parameter_batch shape ({pb.shape[0]},{pb.shape[1]}) doesn't match "
                             f"({trajectory_count},{len(slots)})")
    values, stats, counts, times, kind = _run(
        plan, y0, float(t0), float(t1), abstol=abstol, reltol=reltol, dt=dt, saveat=saveat,
        save_start=save_start, maxiters=maxiters, solver=solver, output=output, output_indices=output_indices,
        integral_weights=integral_weights, slot_indices=slots, parameter_batch=pb,
        trajectory_count=int(trajectory_count), batch=True)
    time_count = int(counts[0])
    width = values.shape[2]
    flat = values[:, :time_count].reshape(-1)
    d = {"solver": solver, "accepted_steps": int(stats[:, 0].sum()), "rejected_steps": int(stats[:, 1].sum()),
         "rhs_calls": int(stats[:, 2].sum()), "trajectory_count": int(trajectory_count)}
    return times[0, :time_count], flat, width, time_count, d


def debug_effective_lindblad_rhs(plan: EffectivePlan, y0, t: float) -> dict[str, Any]:
    """``debug_effective_lindblad_rhs_py`` (``python_api.rs:93-150``) on the GPU."""
    dy, status = plan.device.rhs(np.asarray(y0, dtype=np.float64).reshape(1, -1), np.asarray([float(t)]))
    dy = dy.cpu().numpy()[0]
    st = int(status.cpu().numpy()[0])
    if st:
        raise ValueError("field coordinate outside the operator grid")
    return {"dy": dy, "dy_norm": float(np.sqrt(np.sum(dy * dy))), "dy_max": float(np.max(np.abs(dy))),
            "has_nan": bool(np.isnan(dy).any()), "has_inf": bool(np.isinf(dy).any())}
