"""Terminal stop events (the ``stop_event`` argument of the reference's
solvers, ``centrex_tlf/lindblad/events.py:20-94``).

A stop event is a function g(t, rho) whose zero, or sign change between two
accepted steps, ends a trajectory at the root.  Two kinds exist:

* a population event, g = sum of rho_kk over some states minus a threshold,
  made with :func:`population`;
* an expression event, g = Re of a runtime expression of t and the
  parameters (a ``RuntimeExpression``, a ``Parameter`` or a sympy expression).

:func:`normalize_stop_event` checks an event against a prepared problem and
returns the spec the GPU solver and the oracle take: a dict with ``kind``
"population" (``indices``, ``threshold``) or "runtime_expression" (the
expression compiled to the plan's bytecode over its parameter slots), and a
``name``.
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

_ACCEPTED = "a RuntimeExpression, sympy expression, Parameter, PopulationEvent, or None"


@dataclass(frozen=True)
class PopulationEvent:
    """g = sum of the populations at ``indices`` minus ``threshold``."""

    indices: tuple[int, ...]
    threshold: float = 0.0
    name: str = "population"

    def spec(self, n_states: int) -> dict[str, Any]:
        """The solver's spec, after checking every index against the system size."""
        outside = [k for k in self.indices if not 0 <= k < n_states]
        if outside:
            raise ValueError(f"population event index {outside[0]} out of bounds for {n_states} states")
        return {"kind": "population", "indices": list(self.indices), "threshold": float(self.threshold),
                "name": self.name}


def population(index_or_indices: int | Sequence[int], *, threshold: float = 0.0,
               name: str | None = None) -> PopulationEvent:
    """Event on the summed population of one state or several crossing ``threshold``."""
    picked = tuple(int(k) for k in np.atleast_1d(np.asarray(index_or_indices, dtype=object)).ravel())
    if len(picked) == 0:
        raise ValueError("population event requires at least one index")
    return PopulationEvent(indices=picked, threshold=float(threshold), name="population" if name is None else name)


def _expression_event(event: Any) -> RuntimeExpression:
    match event:
        case RuntimeExpression():
            return event
        case Parameter():
            return event.as_expression()
        case smp.Expr():
            return RuntimeExpression(event, {})
    raise TypeError(f"stop_event must be {_ACCEPTED}")


def normalize_stop_event(stop_event: Any | None, prepared: Any) -> dict[str, Any] | None:
    """The solver's spec of ``stop_event`` for ``prepared`` (None passes through);
    the checks and messages of the reference's ``normalize_stop_event``."""
    if stop_event is None:
        return None
    if isinstance(stop_event, PopulationEvent):
        return stop_event.spec(int(prepared.layout.n))
    runtime = _expression_event(stop_event)
    slot_of = {name: k for k, name in enumerate(prepared.slot_names)}
    unknown = sorted({str(s) for s in runtime.expr.free_symbols} - set(slot_of) - {"t"})
    if unknown:
        raise ValueError(f"stop_event references unknown parameter(s): {', '.join(unknown)}")
    program = compile_expression(runtime.expr, slot_of)
    if program.output_is_tuple:
        raise ValueError("stop_event RuntimeExpression must evaluate to a scalar")
    return {"kind": "runtime_expression", "expression": program, "runtime": runtime, "name": repr(runtime)}
