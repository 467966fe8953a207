"""Option handling shared by ``solve_lindblad`` and the batch/scan drivers.

Restates the reference's Python-side validation and the option translation of
its PyO3 layer (``centrex_tlf/lindblad/solve.py:397-694``,
``centrex_tlf/lindblad/batch.py:41-519``, ``rust/src/lindblad/python_api.rs:422-799``)
and turns a call into one ``DevicePlan.solve`` launch.  Per-trajectory solver
failures come back as the reference's messages (``rust/src/ode/dopri5.rs:170-376``).
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np

from .parameters import Parameter

# execution modes: the GPU (batched persistent kernel) answers for the
# reference's expanded-sparse mode; "gpu_plain" launches one trajectory at a time
# The reference's modes are host RHS algorithms for one equation (dense,
# structured BLAS, structured upper triangle, expanded sparse; rhs.rs:1280-1460):
# on the GPU they all run the expanded sparse program of the plan.
GPU_MODES = {"expanded_sparse": "persistent", "reference": "persistent", "structured": "persistent",
             "structured_upper": "persistent", "gpu": "persistent", "gpu_plain": "plain"}
GPU_SOLVERS = {"dopri5", "tsit5", "fixed_dopri5", "fixed_rk2", "fixed_rk4"}
ADAPTIVE_SOLVERS = {"dopri5", "tsit5"}
FIXED_SOLVERS = {"fixed_dopri5", "fixed_rk2", "fixed_rk4"}
REFERENCE_SOLVERS = {"dopri5", "tsit5", "fixed_dopri5", "fixed_rk2", "fixed_rk4", "scipy_rk45", "scipy_bdf",
                     "scipy_radau", "python_rk45"}
INTEGRAL_OUTPUTS = {"weighted_integral", "photon_integral", "excited_population"}
RATE_OUTPUTS = {"weighted_rate", "photon_rate", "excited_population_rate"}
WEIGHTED_OUTPUTS = INTEGRAL_OUTPUTS | RATE_OUTPUTS
STATUS_MESSAGES = {1: "maxiters", 2: "underflow", 3: "stiff", 4: "bad_span", 5: "bad_support",
                   6: "saveat", 7: "capacity", 8: "bad_dt", 9: "event_complex"}


def time_tol(t: float) -> float:
    """``rust/src/ode/common.rs:11-13``."""
    return 10.0 * np.finfo(np.float64).eps * max(abs(float(t)), 1.0)


def rust_float(x: float) -> str:
    """``format!("{x}")`` of an f64: shortest round-trip digits, positional notation."""
    x = float(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def check_mode_and_solver(execution_mode: str, solver: str) -> str:
    """Launch mode for ``execution_mode`` (``solve.py:548-552``)."""
    if execution_mode not in GPU_MODES:
        raise NotImplementedError(
            "supported execution_mode values are 'reference', 'structured', 'structured_upper', "
            "and 'expanded_sparse' (also 'gpu' and 'gpu_plain')"
        )
    if solver not in GPU_SOLVERS:
        if solver in REFERENCE_SOLVERS:
            raise NotImplementedError(
                f"solver={solver!r} is not available on the GPU; use 'dopri5', 'tsit5', 'fixed_dopri5', "
                "'fixed_rk2' or 'fixed_rk4'"
            )
        raise NotImplementedError(
            "supported solvers are 'dopri5', 'tsit5', 'fixed_dopri5', 'fixed_rk2' and 'fixed_rk4'")
    return GPU_MODES[execution_mode]


def normalize_t_span(t_span: Sequence[float]) -> tuple[float, float]:
    if len(t_span) != 2:
        raise ValueError("t_span must contain exactly two values")
    return (float(t_span[0]), float(t_span[1]))


def normalize_saveat(saveat: Any, t_span: tuple[float, float], save_start: bool) -> np.ndarray | None:
    """``batch.py:47-63`` / ``solve.py:74-90``."""
    if saveat is None:
        return None
    if isinstance(saveat, (float, int, np.floating, np.integer)):
        step = float(saveat)
        if step <= 0:
            raise ValueError("saveat step must be positive")
        values = np.arange(t_span[0], t_span[1] + 0.5 * step, step, dtype=np.float64)
    else:
        values = np.asarray(saveat, dtype=np.float64)
    if not save_start and values.size > 0 and np.isclose(values[0], t_span[0]):
        values = values[1:]
    return values


def _time_tol(t: float) -> float:
    return 10.0 * np.finfo(np.float64).eps * max(abs(t), 1.0)


def build_save_plan(saveat: np.ndarray | None, t0: float, t1: float, save_start: bool) -> np.ndarray | None:
    """``rust/src/ode/common.rs:124-151``: t0 first when saving the start, range checks."""
    if saveat is None:
        return None
    times = [t0] if save_start else []
    for t in np.asarray(saveat, dtype=np.float64).reshape(-1):
        t = float(t)
        if save_start and abs(t - t0) <= _time_tol(t0):
            continue
        if not np.isfinite(t):
            raise ValueError("saveat contains non-finite value")
        if t < t0 - _time_tol(t0) or t > t1 + _time_tol(t1):
            raise ValueError(f"saveat value {rust_float(t)} is outside [{rust_float(t0)}, {rust_float(t1)}]")
        times.append(t)
    return np.asarray(times, dtype=np.float64)


def parameter_slot_name(slot: Any) -> str:
    if isinstance(slot, Parameter):
        return slot.name
    return str(slot)


def parameter_slot_indices(prepared: Any, parameter_slots: Sequence[Any] | None) -> list[int]:
    """``batch.py:94-116``."""
    if parameter_slots is None:
        return []
    slot_names = list(prepared.parameter_graph["slot_names"])
    base_count = len(prepared.parameter_graph["base_values"])
    indices = []
    for slot in parameter_slots:
        name = parameter_slot_name(slot)
        try:
            index = slot_names.index(name)
        except ValueError as exc:
            raise ValueError(f"unknown parameter slot {name!r}") from exc
        if index >= base_count:
            raise ValueError(
                f"parameter slot {name!r} is a compound parameter; only base parameters can be scanned"
            )
        indices.append(index)
    return indices


def check_outputs(output: str, output_indices: Any, output_when: str, integral_weights: Any,
                  saveat_values: np.ndarray | None, *, allow_full: bool, batch: bool) -> None:
    """Validation of ``batch.py:170-207`` / ``solve.py:548-592`` (same messages)."""
    allowed = {"populations", "selected", *WEIGHTED_OUTPUTS} | ({"full"} if allow_full else set())
    if output not in allowed:
        if allow_full:
            raise NotImplementedError(
                "output must be 'full', 'populations', 'selected', 'weighted_integral', "
                "'photon_integral', 'excited_population', 'weighted_rate', 'photon_rate', "
                "or 'excited_population_rate'"
            )
        raise NotImplementedError(
            "batch solving currently supports output='populations', 'selected', "
            "'weighted_integral', 'photon_integral', 'excited_population', "
            "'weighted_rate', 'photon_rate', or 'excited_population_rate'"
        )
    if output_when not in {"final", "saveat"}:
        if batch:
            raise ValueError("output_when must be 'final' or 'saveat'")
        raise NotImplementedError("output_when must be 'saveat' or 'final'")
    if output == "selected" and output_indices is None:
        raise ValueError("output='selected' requires output_indices")
    if output != "selected" and output_indices is not None:
        raise ValueError("output_indices is only valid with output='selected'")
    if output in WEIGHTED_OUTPUTS and integral_weights is None:
        raise ValueError(f"output={output!r} requires integral_weights")
    if output not in WEIGHTED_OUTPUTS and integral_weights is not None:
        raise ValueError("integral_weights are only valid with weighted output modes")
    if output in RATE_OUTPUTS and output_when != "saveat":
        raise ValueError(f"output={output!r} requires output_when='saveat'")
    if output in RATE_OUTPUTS and saveat_values is None:
        raise ValueError(f"output={output!r} requires explicit saveat values")


def trajectory_message(status: int, fail_t: float, *, maxiters: int, n_plan: int, solver: str = "dopri5") -> str:
    """The reference's error text for a failed trajectory (dopri5.rs:226-231,
    fixed.rs:154-156, 185-190: the fixed-step loop counts steps from 0)."""
    if status == 1:
        steps = maxiters if solver in FIXED_SOLVERS else maxiters + 1
        return f"Stopped at x={rust_float(fail_t)}. Need more than {steps} steps."
    if status == 2:
        return f"Stopped at x={rust_float(fail_t)}. Step size underflow."
    if status == 3:
        return f"Stiff at x={rust_float(fail_t)}."
    if status == 4:
        return "only forward integration"
    if status == 6:
        return f"ended before all saveat: {int(fail_t)} of {n_plan}"
    if status == 5:
        return "initial state is nonzero outside the declared support"
    if status == 8:
        return "fixed-step solvers require positive finite dt"
    if status == 9:
        return "event expression evaluated to a complex value"
    return f"solver failed with status {status}"


@dataclass
class DeviceRun:
    """Raw result of one launch (host arrays)."""

    values: np.ndarray          # (B, width) or (B, capacity, width)
    stats: np.ndarray           # (B, 4)
    fail_t: np.ndarray          # (B,)
    counts: np.ndarray | None   # (B,) rows pushed, trace mode
    times: np.ndarray | None    # (B, capacity), trace mode with times
    event_t: np.ndarray | None = None  # (B,) stop-event root times, NaN where none fired


def run_device(prepared: Any, y0: np.ndarray, t0: float, t1: float, *, solver: str, mode: str,
               abstol: float, reltol: float, dt: float, maxiters: int, save_start: bool,
               kernel_output: str, selected: Sequence[tuple[int, int]] | None,
               weights: Sequence[tuple[int, float]] | None, slot_indices: Sequence[int],
               parameter_values: np.ndarray | None, grid_axes: Sequence[np.ndarray] | None,
               save_plan: np.ndarray | None, trace: bool, capacity: int, with_times: bool,
               stop_event: dict[str, Any] | None = None) -> DeviceRun:
    """One ``lb_solve`` launch; y0 is (dim,) shared or (B, dim)."""
    plan = prepared.device_plan()
    kw: dict[str, Any] = dict(
        solver=solver, abstol=float(abstol), reltol=float(reltol), dt=float(dt), maxiters=int(maxiters),
        save_start=bool(save_start), output_kind=kernel_output, selected=selected, weights=weights,
        mode=mode, save_plan=save_plan, trace=trace, capacity=int(capacity), with_times=with_times,
        stop_event=stop_event,
    )
    if grid_axes is not None:
        kw.update(override_slots=list(slot_indices), grid_axes=grid_axes)
    elif slot_indices:
        kw.update(override_slots=list(slot_indices), override_values=parameter_values)
    elif np.ndim(y0) == 1:
        kw.update(n_traj=1)
    out = plan.solve(np.ascontiguousarray(y0, dtype=np.float64), t0, t1, **kw)
    event_t = out[-1].cpu().numpy() if stop_event is not None else None
    if trace:
        values, stats, fail_t, _, counts, times = out[:6]
        return DeviceRun(values.cpu().numpy(), stats.cpu().numpy(), fail_t.cpu().numpy(),
                         counts.cpu().numpy(), None if times is None else times.cpu().numpy(), event_t)
    values, stats, fail_t = out[:3]
    return DeviceRun(values.cpu().numpy(), stats.cpu().numpy(), fail_t.cpu().numpy(), None, None, event_t)


def kernel_output_kind(output: str) -> str:
    if output in INTEGRAL_OUTPUTS:
        return "integral"
    if output in RATE_OUTPUTS:
        return "rate"
    return output


def width_of(prepared: Any, output: str, output_indices: Any) -> int:
    if output == "populations":
        return prepared.layout.n
    if output == "full":
        return prepared.layout.packed_len
    if output == "selected":
        return len(output_indices)
    return 1


def stats_dict(solver: str, stats: np.ndarray, *, saved_points: int, elapsed: float,
               batch: bool, trajectory_count: int, event_t: np.ndarray | None = None,
               end_times: np.ndarray | None = None) -> dict[str, Any]:
    """Solver stats with the reference's keys (``python_api.rs:516-527, 654-667``,
    ``batch.py:127-140``): per-trajectory event flags and times (the root, or
    the last output time ``end_times`` where no event fired)."""
    fired = np.zeros(trajectory_count, dtype=bool) if event_t is None else np.isfinite(event_t)
    d: dict[str, Any] = {
        "solver": solver,
        "accepted_steps": int(stats[:, 0].sum()),
        "rejected_steps": int(stats[:, 1].sum()),
        "rhs_calls": int(stats[:, 2].sum()),
        "event_triggered": bool(fired.any()),
    }
    if batch:
        d["event_triggered_by_trajectory"] = [bool(f) for f in fired]
        ends = np.zeros(trajectory_count) if end_times is None else np.broadcast_to(end_times, (trajectory_count,))
        d["event_times"] = [float(e) if f else float(x) for f, e, x in
                            zip(fired, event_t if event_t is not None else ends, ends)]
        d["event_count"] = int(fired.sum())
    elif fired.any():
        d["event_time"] = float(event_t[0])
        d["event_index"] = 0
    d["function_evaluations"] = d["rhs_calls"]
    d["saved_points"] = saved_points
    d["rhs_seconds"] = elapsed
    d["total_seconds"] = elapsed
    return d
