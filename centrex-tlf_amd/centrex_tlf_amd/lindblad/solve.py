"""``solve_lindblad``: one trajectory on the GPU, with the reference's arguments and results.

Mirrors ``centrex_tlf/lindblad/solve.py:397-694`` for the Dopri5/Tsit5
expanded-sparse path: ``output`` in full / populations / selected / weighted
integral / weighted rate, ``output_when`` "saveat" (a row at every pushed point:
t0 when ``save_start`` and every accepted step, or the ``saveat`` times through
the solver's dense output) or "final".  Failures raise ``ValueError`` with the
reference solver's message.
"""

from __future__ import annotations

import time
from collections.abc import Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np
import numpy.typing as npt

from . import _driver as D
from .events import normalize_stop_event
from .plan_static import PreparedLindbladProblem, prepare_lindblad_problem

__all__ = ["LindbladObservableResult", "LindbladResult", "solve_lindblad"]

_FIRST_TRACE_CAPACITY = 4096


@dataclass
class LindbladResult:
    """Packed states at the saved times (``solve.py:26-41``)."""

    t: npt.NDArray[np.float64]
    packed_y: npt.NDArray[np.float64]
    layout: Any
    solver_stats: dict[str, Any] | None = None

    def density_matrices(self) -> npt.NDArray[np.complex128]:
        return self.layout.unpack_batch(self.packed_y)

    def populations(self) -> npt.NDArray[np.float64]:
        return np.ascontiguousarray(self.packed_y[:, : self.layout.n])


@dataclass
class LindbladObservableResult:
    """Reduced outputs (``solve.py:65-71``)."""

    t: npt.NDArray[np.float64]
    values: npt.NDArray[Any]
    output: str
    output_indices: Any | None = None
    solver_stats: dict[str, Any] | None = None


def solve_lindblad(
    prepared_or_obe_system: PreparedLindbladProblem | Any,
    rho0: npt.NDArray[np.complex128],
    t_span: Sequence[float],
    *,
    parameters: Any | None = None,
    backend: str = "gpu",
    solver: str | None = None,
    abstol: float = 1e-7,
    reltol: float = 1e-4,
    dt: float = 1e-8,
    saveat: None | float | Sequence[float] | npt.NDArray[np.floating] = None,
    save_start: bool = True,
    maxiters: int = 100_000,
    execution_mode: str = "expanded_sparse",
    use_split_input_rhs: bool = True,
    jacobian: str = "exact",
    jacobian_format: str = "auto",
    collect_stats: bool = False,
    output: str = "full",
    output_indices: Sequence[tuple[int, int]] | None = None,
    output_when: str = "saveat",
    dense_output: bool = True,
    integral_weights: Sequence[tuple[int, float]] | None = None,
    stop_event: Any | None = None,
) -> LindbladResult | LindbladObservableResult:
    """Integrate one density matrix on the GPU (see the module docstring).

    ``use_split_input_rhs``, ``jacobian`` and ``jacobian_format`` are accepted
    for signature compatibility; the GPU kernel chooses its own RHS layout and
    the explicit solvers use no Jacobian.
    """
    if solver is None:
        solver = "dopri5"
    if backend not in {"gpu", "rust"}:
        raise NotImplementedError(f"backend={backend!r} is not available; this solver runs on the GPU")
    launch_mode = D.check_mode_and_solver(execution_mode, solver)
    t_span_tuple = D.normalize_t_span(t_span)
    saveat_values = D.normalize_saveat(saveat, t_span_tuple, save_start)
    D.check_outputs(output, output_indices, output_when, integral_weights, saveat_values,
                    allow_full=True, batch=False)
    if not dense_output and output_when == "saveat" and saveat_values is not None:
        interior = saveat_values[~np.isclose(saveat_values, t_span_tuple[0])
                                 & ~np.isclose(saveat_values, t_span_tuple[1])]
        if interior.size:
            raise ValueError("dense_output=False cannot be used with interior saveat points")
    if isinstance(prepared_or_obe_system, PreparedLindbladProblem):
        prepared = prepared_or_obe_system
    else:
        if parameters is None:
            raise TypeError("parameters are required when solving from an OBESystem")
        prepared = prepare_lindblad_problem(prepared_or_obe_system, parameters, backend="gpu")
    prepared.check_execution_mode(execution_mode)
    packed_rho0 = prepared.layout.pack(np.asarray(rho0, dtype=np.complex128))
    t0, t1 = t_span_tuple
    event_spec = normalize_stop_event(stop_event, prepared)

    # option translation of python_api.rs:422-528 / solve.py:421-429
    integral = output in D.INTEGRAL_OUTPUTS
    final = output_when == "final"
    kernel_output = D.kernel_output_kind(output)
    width = D.width_of(prepared, output, output_indices)
    eff_save_start = save_start and not (final and not integral)
    save_plan = None if (final and not integral) else D.build_save_plan(saveat_values, t0, t1, eff_save_start)
    trace = not final
    if trace and save_plan is not None:
        capacity = max(1, save_plan.size)
    else:
        capacity = min(int(maxiters) + 2, _FIRST_TRACE_CAPACITY)
    selected = None if output_indices is None else [(int(i), int(j)) for i, j in output_indices]
    start = time.perf_counter()
    while True:
        run = D.run_device(
            prepared, packed_rho0, t0, t1, solver=solver, mode=launch_mode, abstol=abstol, reltol=reltol,
            dt=dt, maxiters=maxiters, save_start=eff_save_start, kernel_output=kernel_output,
            selected=selected, weights=integral_weights, slot_indices=[], parameter_values=None,
            grid_axes=None, save_plan=save_plan, trace=trace, capacity=capacity,
            with_times=trace and (save_plan is None or event_spec is not None), stop_event=event_spec)
        status = int(run.stats[0, 3])
        if status == 7:  # more accepted steps than rows: the counts are exact now
            capacity = int(run.counts[0])
            continue
        break
    elapsed = time.perf_counter() - start
    if status != 0:
        raise ValueError(D.trajectory_message(status, float(run.fail_t[0]), maxiters=maxiters,
                                              n_plan=0 if save_plan is None else save_plan.size, solver=solver))
    if trace:
        count = int(run.counts[0])
        rows = run.values[0, :count]
        times = run.times[0, :count] if run.times is not None else save_plan[:count]
    else:
        rows = run.values[0]
        times = np.asarray([t1], dtype=np.float64)
    if integral and final:
        times = np.asarray([t1 if save_plan is None else (save_plan[-1] if save_plan.size else t1)])
    fired = event_spec is not None and bool(np.isfinite(run.event_t[0]))
    if fired and final:
        # the root is the last pushed point -- unless a state output's save
        # point t1 lies within time_tol of it (dopri5.rs:318-352)
        te = float(run.event_t[0])
        times = np.asarray([t1 if (not integral and t1 <= te + D.time_tol(te)) else te])
    stats = (D.stats_dict(solver, run.stats, saved_points=int(times.size), elapsed=elapsed, batch=False,
                          trajectory_count=1, event_t=run.event_t) if collect_stats else None)
    if stats is not None and fired:
        stats["event_name"] = event_spec.get("name", "event")
    if output == "full":
        packed = np.asarray(rows, dtype=np.float64).reshape(times.size, width)
        return LindbladResult(t=np.asarray(times, dtype=np.float64), packed_y=packed, layout=prepared.layout,
                              solver_stats=stats)
    if output in D.WEIGHTED_OUTPUTS:
        vals = np.asarray(rows, dtype=np.float64).reshape(-1)
        if output_when == "saveat":
            vals = vals.reshape(times.size, 1)
        return LindbladObservableResult(t=np.asarray(times, dtype=np.float64), values=vals, output=output,
                                        output_indices=None, solver_stats=stats)
    if output == "selected":
        vals = np.asarray(rows, dtype=np.float64).reshape(-1, 2 * width)
        vals = (vals[:, 0::2] + 1j * vals[:, 1::2]).astype(np.complex128)
    else:
        vals = np.asarray(rows, dtype=np.float64).reshape(-1, width)
    if final:
        vals = vals.reshape(width)
    return LindbladObservableResult(t=np.asarray(times, dtype=np.float64), values=vals, output=output,
                                    output_indices=None if output_indices is None else list(output_indices),
                                    solver_stats=stats)
