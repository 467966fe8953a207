"""Terminal stop events (``centrex_tlf/lindblad/events.py:20-94``).

A stop event is a function g(t, y) whose sign change (or zero) between two
accepted steps ends a trajectory at the root: the sum of some populations
minus a threshold (``population(...)``), or a real runtime expression of t and
the parameters (a ``RuntimeExpression``, ``Parameter`` or sympy expression).
``normalize_stop_event`` checks one against a prepared problem and returns the
spec the GPU solver takes: the population indices and threshold, or the
expression compiled to the plan's bytecode over the plan's parameter slots.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np
import sympy as smp

from .ir import compile_expression
from .parameters import Parameter, RuntimeExpression

__all__ = ["PopulationEvent", "normalize_stop_event", "population"]


@dataclass(frozen=True)
class PopulationEvent:
    indices: tuple[int, ...]
    threshold: float = 0.0
    name: str = "population"


def population(index_or_indices: int | Sequence[int], *, threshold: float = 0.0,
               name: str | None = None) -> PopulationEvent:
    """Event on the sum of the populations at ``index_or_indices`` crossing ``threshold``."""
    if isinstance(index_or_indices, int | np.integer):
        indices = (int(index_or_indices),)
    else:
        indices = tuple(int(index) for index in index_or_indices)
    if not indices:
        raise ValueError("population event requires at least one index")
    return PopulationEvent(indices=indices, threshold=float(threshold),
                           name=name if name is not None else "population")


def _as_runtime_expression(stop_event: Any) -> RuntimeExpression | None:
    if isinstance(stop_event, RuntimeExpression):
        return stop_event
    if isinstance(stop_event, Parameter):
        return stop_event.as_expression()
    if isinstance(stop_event, smp.Expr):
        return RuntimeExpression(stop_event, {})
    return None


def normalize_stop_event(stop_event: Any | None, prepared: Any) -> dict[str, Any] | None:
    """The event spec for the solver (messages as ``events.py:57-94``)."""
    if stop_event is None:
        return None
    n_states = int(prepared.layout.n)
    if isinstance(stop_event, PopulationEvent):
        indices = [int(index) for index in stop_event.indices]
        for index in indices:
            if index < 0 or index >= n_states:
                raise ValueError(f"population event index {index} out of bounds for {n_states} states")
        return {"kind": "population", "indices": indices, "threshold": float(stop_event.threshold),
                "name": stop_event.name}
    runtime = _as_runtime_expression(stop_event)
    if runtime is None:
        raise TypeError("stop_event must be a RuntimeExpression, sympy expression, Parameter, "
                        "PopulationEvent, or None")
    slot_index_by_name = {name: idx for idx, name in enumerate(prepared.slot_names)}
    missing = sorted(str(s) for s in runtime.expr.free_symbols
                     if str(s) != "t" and str(s) not in slot_index_by_name)
    if missing:
        raise ValueError(f"stop_event references unknown parameter(s): {', '.join(missing)}")
    compiled = compile_expression(runtime.expr, slot_index_by_name)
    if compiled.output_is_tuple:
        raise ValueError("stop_event RuntimeExpression must evaluate to a scalar")
    return {"kind": "runtime_expression", "expression": compiled, "runtime": runtime, "name": repr(runtime)}
