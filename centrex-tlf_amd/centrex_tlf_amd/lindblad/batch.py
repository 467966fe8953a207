"""Batch and scan drivers on the GPU with the reference's arguments and result layouts.

Mirrors ``centrex_tlf/lindblad/batch.py:143-519``: ``solve_lindblad_batch``
(one initial state and optional parameter row per trajectory),
``initial_condition_scan``, ``parameter_scan`` and ``grid_scan`` (Cartesian
product of base-parameter axes, last axis fastest, one shared initial state).
All trajectories of a call run in one launch of the persistent kernel
(``execution_mode="expanded_sparse"`` or ``"gpu"``) or one launch each
(``"gpu_plain"``).  Values are ``(B, width)`` for ``output_when="final"`` and
``(B, T, width)`` for ``"saveat"``; a failing trajectory raises
``ValueError("trajectory {i}: ...")`` like the reference's batch driver.
"""

from __future__ import annotations

import time
from collections.abc import Mapping, Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import numpy.typing as npt

from . import _driver as D
from .events import normalize_stop_event
from .plan_static import PreparedLindbladProblem

__all__ = [
    "LindbladBatchResult",
    "grid_scan",
    "initial_condition_scan",
    "parameter_scan",
    "solve_lindblad_batch",
]


@dataclass
class LindbladBatchResult:
    """``batch.py:29-39``."""

    t: npt.NDArray[np.float64]
    values: npt.NDArray[Any]
    output: str
    output_indices: list[tuple[int, int]] | None
    trajectory_count: int
    parameter_slots: list[str] | None = None
    parameter_values: npt.NDArray[np.complex128] | None = None
    solver_stats: dict[str, Any] | None = None
    metadata: dict[str, Any] = field(default_factory=dict)


def _pack_rho0_batch(prepared: PreparedLindbladProblem, rho0_batch: np.ndarray) -> np.ndarray:
    """``batch.py:66-91`` (vectorised packing)."""
    batch = np.asarray(rho0_batch)
    if batch.ndim == 2:
        packed = np.asarray(batch, dtype=np.float64)
        if packed.shape[1] != prepared.layout.packed_len:
            raise ValueError(
                "2D rho0_batch must have shape (n_trajectories, packed_len); "
                f"expected packed_len={prepared.layout.packed_len}, got {packed.shape[1]}"
            )
        return np.ascontiguousarray(packed)
    if batch.ndim != 3:
        raise ValueError(
            "rho0_batch must have shape (n_trajectories, packed_len) or "
            "(n_trajectories, n_states, n_states)"
        )
    if batch.shape[1:] != (prepared.layout.n, prepared.layout.n):
        raise ValueError(
            "matrix rho0_batch must have shape "
            f"(n_trajectories, {prepared.layout.n}, {prepared.layout.n})"
        )
    return np.ascontiguousarray(prepared.layout.pack_batch(np.asarray(batch, dtype=np.complex128)))


def _pack_rho0(prepared: PreparedLindbladProblem, rho0: np.ndarray) -> np.ndarray:
    rho = np.asarray(rho0)
    if rho.ndim == 1:
        packed = np.asarray(rho, dtype=np.float64)
        if packed.size != prepared.layout.packed_len:
            raise ValueError(f"packed rho0 must have length {prepared.layout.packed_len}, got {packed.size}")
        return packed
    if rho.ndim == 2:
        return prepared.layout.pack(np.asarray(rho, dtype=np.complex128))
    raise ValueError("rho0 must be a packed vector or density matrix")


def _common_checks(prepared: PreparedLindbladProblem, *, solver: str, execution_mode: str, output: str,
                   output_indices: Any, output_when: str, integral_weights: Any,
                   saveat_values: np.ndarray | None, stop_event: Any, threads: Any, dense_output: bool,
                   name: str) -> str:
    """``batch.py:170-207`` / ``batch.py:402-443`` (same checks and messages)."""
    prepared.check_execution_mode(execution_mode)
    mode = D.check_mode_and_solver(execution_mode, solver)
    D.check_outputs(output, output_indices, output_when, integral_weights, saveat_values,
                    allow_full=False, batch=True)
    if output_when == "saveat" and saveat_values is None:
        raise ValueError(f"{name} output_when='saveat' requires explicit saveat values")
    if stop_event is not None and output_when != "final":
        raise ValueError(f"{name} stop_event is only supported with output_when='final'")
    if not dense_output and output_when == "saveat" and saveat_values is not None:
        raise ValueError("dense_output=False is only supported with output_when='final'")
    if threads is not None and threads <= 0:
        raise ValueError("threads must be positive when provided")
    return mode


def _run(prepared: PreparedLindbladProblem, y0: np.ndarray, t_span: tuple[float, float], *, mode: str,
         solver: str, abstol: float, reltol: float, dt: float, saveat_values: np.ndarray | None,
         save_start: bool, maxiters: int, output: str, output_indices: Any, output_when: str,
         integral_weights: Any, slot_indices: list[int], parameter_values: np.ndarray | None,
         grid_axes: list[np.ndarray] | None, trajectory_count: int, collect_stats: bool,
         event_spec: dict[str, Any] | None = None):
    t0, t1 = t_span
    integral = output in D.INTEGRAL_OUTPUTS
    final = output_when == "final"
    # python_api.rs:589-617: a final non-integral output is the state at t1
    eff_save_start = (not final or integral) and save_start
    save_plan = None if (final and not integral) else D.build_save_plan(saveat_values, t0, t1, eff_save_start)
    trace = not final
    capacity = max(1, save_plan.size) if (trace and save_plan is not None) else 1
    width = D.width_of(prepared, output, output_indices)
    selected = None if output_indices is None else [(int(i), int(j)) for i, j in output_indices]
    start = time.perf_counter()
    run = D.run_device(prepared, y0, t0, t1, solver=solver, mode=mode, abstol=abstol, reltol=reltol, dt=dt,
                       maxiters=maxiters, save_start=eff_save_start, kernel_output=D.kernel_output_kind(output),
                       selected=selected, weights=integral_weights, slot_indices=slot_indices,
                       parameter_values=parameter_values, grid_axes=grid_axes, save_plan=save_plan,
                       trace=trace, capacity=capacity, with_times=False, stop_event=event_spec)
    elapsed = time.perf_counter() - start
    bad = np.flatnonzero(run.stats[:, 3] != 0)
    if bad.size:
        i = int(bad[0])
        msg = D.trajectory_message(int(run.stats[i, 3]), float(run.fail_t[i]), maxiters=maxiters,
                                   n_plan=0 if save_plan is None else save_plan.size, solver=solver)
        raise ValueError(f"trajectory {i}: {msg}")
    if trace:
        time_count = int(run.counts[0])
        times = save_plan[:time_count]
        raw = run.values[:, :time_count]
    else:
        time_count = 1
        times = np.asarray([t1 if (save_plan is None or not save_plan.size) else save_plan[-1]]
                           if integral else [t1], dtype=np.float64)
        raw = run.values
    if output == "selected":
        raw = raw.reshape(raw.shape[:-1] + (width, 2))
        values = (raw[..., 0] + 1j * raw[..., 1]).astype(np.complex128)
    else:
        values = np.asarray(raw, dtype=np.float64)
    values = values.reshape((trajectory_count, width) if final else (trajectory_count, time_count, width))
    end_time = float(times[-1]) if len(times) else t1
    events = None
    if event_spec is not None:
        # ode_batch.rs:305-340: t holds each trajectory's root, or its last output time
        fired = np.isfinite(run.event_t)
        times = np.where(fired, run.event_t, end_time)
        events = {"event_triggered": fired, "event_times": np.asarray(times, dtype=np.float64)}
    stats = None
    if collect_stats:
        stats = D.stats_dict(solver, run.stats, saved_points=time_count, elapsed=elapsed, batch=True,
                             trajectory_count=trajectory_count, event_t=run.event_t,
                             end_times=np.full(trajectory_count, end_time))
    return np.asarray(times, dtype=np.float64), values, stats, events


def solve_lindblad_batch(
    prepared: PreparedLindbladProblem,
    rho0_batch: npt.NDArray[np.complex128] | npt.NDArray[np.float64],
    t_span: Sequence[float],
    *,
    parameter_batch: npt.NDArray[np.complex128] | None = None,
    parameter_slots: Sequence[Any] | None = None,
    solver: str = "dopri5",
    execution_mode: str = "expanded_sparse",
    use_split_input_rhs: bool = True,
    abstol: float = 1e-7,
    reltol: float = 1e-4,
    dt: float = 1e-8,
    saveat: None | float | Sequence[float] | npt.NDArray[np.floating] = None,
    save_start: bool = True,
    maxiters: int = 100_000,
    collect_stats: bool = False,
    output: str = "populations",
    output_indices: Sequence[tuple[int, int]] | None = None,
    output_when: str = "final",
    integral_weights: Sequence[tuple[int, float]] | None = None,
    dense_output: bool = True,
    parallel: bool = True,
    threads: int | None = None,
    metadata: Mapping[str, Any] | None = None,
    stop_event: Any | None = None,
) -> LindbladBatchResult:
    """Solve a batch of trajectories (``batch.py:143-313``); ``parallel``/``threads``/
    ``use_split_input_rhs`` are accepted and have no effect on the GPU."""
    t_span_tuple = D.normalize_t_span(t_span)
    saveat_values = D.normalize_saveat(saveat, t_span_tuple, save_start)
    mode = _common_checks(prepared, solver=solver, execution_mode=execution_mode, output=output,
                          output_indices=output_indices, output_when=output_when,
                          integral_weights=integral_weights, saveat_values=saveat_values,
                          stop_event=stop_event, threads=threads, dense_output=dense_output,
                          name="batch")
    packed_batch = _pack_rho0_batch(prepared, np.asarray(rho0_batch))
    trajectory_count = packed_batch.shape[0]
    if trajectory_count == 0:
        raise ValueError("rho0_batch must contain at least one trajectory")
    slot_indices = D.parameter_slot_indices(prepared, parameter_slots)
    parameter_values = None
    if parameter_batch is not None:
        if parameter_slots is None:
            raise ValueError("parameter_slots are required when parameter_batch is provided")
        parameter_values = np.ascontiguousarray(parameter_batch, dtype=np.complex128)
        if parameter_values.shape != (trajectory_count, len(slot_indices)):
            raise ValueError(
                "parameter_batch must have shape "
                f"({trajectory_count}, {len(slot_indices)}), got {parameter_values.shape}"
            )
    elif parameter_slots is not None:
        raise ValueError("parameter_slots were provided without parameter_batch")
    event_spec = normalize_stop_event(stop_event, prepared)
    times, values, stats, events = _run(
        prepared, packed_batch, t_span_tuple, mode=mode, solver=solver, abstol=abstol, reltol=reltol, dt=dt,
        saveat_values=saveat_values, save_start=save_start, maxiters=maxiters, output=output,
        output_indices=output_indices, output_when=output_when, integral_weights=integral_weights,
        slot_indices=slot_indices, parameter_values=parameter_values, grid_axes=None,
        trajectory_count=trajectory_count, collect_stats=collect_stats, event_spec=event_spec)
    return LindbladBatchResult(
        t=times,
        values=values,
        output=output,
        output_indices=None if output_indices is None else list(output_indices),
        trajectory_count=trajectory_count,
        parameter_slots=None if parameter_slots is None else [D.parameter_slot_name(s) for s in parameter_slots],
        parameter_values=parameter_values,
        solver_stats=stats,
        metadata={**({} if metadata is None else dict(metadata)), **(events or {})},
    )


def initial_condition_scan(
    prepared: PreparedLindbladProblem,
    rho0_batch: npt.NDArray[np.complex128] | npt.NDArray[np.float64],
    t_span: Sequence[float],
    **kwargs: Any,
) -> LindbladBatchResult:
    """``batch.py:316-322``."""
    return solve_lindblad_batch(prepared, rho0_batch, t_span, **kwargs)


def parameter_scan(
    prepared: PreparedLindbladProblem,
    rho0: npt.NDArray[np.complex128] | npt.NDArray[np.float64],
    t_span: Sequence[float],
    *,
    parameter_slots: Sequence[Any],
    parameter_batch: npt.NDArray[np.complex128],
    **kwargs: Any,
) -> LindbladBatchResult:
    """``batch.py:325-350``: one initial state, one parameter row per trajectory."""
    values = np.asarray(parameter_batch, dtype=np.complex128)
    if values.ndim != 2:
        raise ValueError("parameter_batch must be 2D")
    packed = _pack_rho0(prepared, rho0)
    rho0_batch = np.broadcast_to(packed.reshape(1, -1), (values.shape[0], packed.size))
    return solve_lindblad_batch(prepared, rho0_batch, t_span, parameter_slots=parameter_slots,
                                parameter_batch=values, **kwargs)


# grid_scan's keywords: the batch call's, with the same defaults; a conversion
# (or None) applied to the given or default value.  use_split_input_rhs and
# parallel are accepted and have no effect (one GPU kernel family serves every
# system size, and trajectories always run in parallel).
_GRID_OPTIONS: tuple[tuple[str, Any, Any], ...] = (
    ("save_start", True, bool), ("saveat", None, None), ("solver", "dopri5", None),
    ("execution_mode", "expanded_sparse", None), ("use_split_input_rhs", True, None),
    ("abstol", 1e-7, float), ("reltol", 1e-4, float), ("dt", 1e-8, float), ("maxiters", 100_000, int),
    ("collect_stats", False, bool), ("output", "populations", None), ("output_indices", None, None),
    ("output_when", "final", None), ("integral_weights", None, None), ("dense_output", True, bool),
    ("parallel", True, None), ("threads", None, None), ("metadata", None, None), ("stop_event", None, None),
)


def _grid_options(given: Mapping[str, Any]) -> dict[str, Any]:
    known = {name for name, _, _ in _GRID_OPTIONS}
    extra = sorted(set(given) - known)
    if extra:
        raise TypeError(f"unexpected grid_scan keyword argument(s): {', '.join(extra)}")
    options = {}
    for name, default, convert in _GRID_OPTIONS:
        value = given.get(name, default)
        options[name] = value if convert is None else convert(value)
    return options


def grid_scan(
    prepared: PreparedLindbladProblem,
    rho0: npt.NDArray[np.complex128] | npt.NDArray[np.float64],
    t_span: Sequence[float],
    *,
    scan: Mapping[Any, Sequence[complex] | npt.NDArray[np.complexfloating]],
    **kwargs: Any,
) -> LindbladBatchResult:
    """``batch.py:353-519``: Cartesian grid over base parameters, last axis fastest,
    the axes uploaded once and decoded per trajectory on the device."""
    if not scan:
        raise ValueError("scan must contain at least one parameter")
    parameter_slots = list(scan)
    parameter_slot_names = [D.parameter_slot_name(s) for s in parameter_slots]
    axes = [np.asarray(values, dtype=np.complex128).reshape(-1) for values in scan.values()]
    if any(axis.size == 0 for axis in axes):
        raise ValueError("scan axes must be non-empty")
    slot_indices = D.parameter_slot_indices(prepared, parameter_slots)
    packed = _pack_rho0(prepared, rho0)
    t_span_tuple = D.normalize_t_span(t_span)
    o = _grid_options(kwargs)
    save_start = o["save_start"]
    saveat_values = D.normalize_saveat(o["saveat"], t_span_tuple, save_start)
    solver, execution_mode, output = o["solver"], o["execution_mode"], o["output"]
    abstol, reltol, dt, maxiters = o["abstol"], o["reltol"], o["dt"], o["maxiters"]
    output_indices, output_when, integral_weights = o["output_indices"], o["output_when"], o["integral_weights"]
    collect_stats, dense_output, threads = o["collect_stats"], o["dense_output"], o["threads"]
    metadata, stop_event = o["metadata"], o["stop_event"]
    mode = _common_checks(prepared, solver=solver, execution_mode=execution_mode, output=output,
                          output_indices=output_indices, output_when=output_when,
                          integral_weights=integral_weights, saveat_values=saveat_values,
                          stop_event=stop_event, threads=threads, dense_output=dense_output,
                          name="grid_scan")
    trajectory_count = int(np.prod([a.size for a in axes], dtype=np.int64))
    event_spec = normalize_stop_event(stop_event, prepared)
    times, values, stats, events = _run(
        prepared, packed, t_span_tuple, mode=mode, solver=solver, abstol=abstol, reltol=reltol, dt=dt,
        saveat_values=saveat_values, save_start=save_start, maxiters=maxiters, output=output,
        output_indices=output_indices, output_when=output_when, integral_weights=integral_weights,
        slot_indices=slot_indices, parameter_values=None, grid_axes=axes,
        trajectory_count=trajectory_count, collect_stats=collect_stats, event_spec=event_spec)
    result = LindbladBatchResult(
        t=times,
        values=values,
        output=output,
        output_indices=None if output_indices is None else list(output_indices),
        trajectory_count=trajectory_count,
        parameter_slots=parameter_slot_names,
        parameter_values=None,
        solver_stats=stats,
        metadata={**({} if metadata is None else dict(metadata)), **(events or {})},
    )
    result.metadata.update({
        "scan_kind": "grid",
        "grid_shape": tuple(int(axis.size) for axis in axes),
        "grid_axes": dict(zip(parameter_slot_names, axes)),
        "compact_grid": True,
    })
    return result
