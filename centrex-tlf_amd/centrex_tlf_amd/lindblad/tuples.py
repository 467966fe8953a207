"""Tuple-valued parameters, lowered to scalar programs for the device.

The reference evaluates tuples at run time (``RuntimeValue::Tuple``,
``rust/src/lindblad/eval.rs:88-133, 500-656``): arithmetic, NEG and CONJ
broadcast over tuples elementwise, builtins, comparisons and the scalar
helpers reject them, and four helpers take tuple arguments
(``eval.rs:250-451``): ``multipass_2d_intensity`` (8), ``multipass_2d_rabi``
(10), ``linear_interp`` (14) and ``pchip_interp`` (16).

Tuple lengths are fixed when a problem is prepared (only base parameters hold
tuples, and a batch overrides scalars), so the structure is resolved here,
once, and the device machine stays scalar:

* a read of a base tuple becomes its elements, as constants;
* a tuple-valued compound becomes one scalar compound per element, in new
  slots after the plan's slots (its own slot is left unused);
* an elementwise operation is written out per element (the same operation on
  the same operands, so each element rounds as the reference's);
* a tuple argument of a helper is passed as its elements, consecutively:
  argc = 1 + 2L for 14, 16 and 17 (x, grid, values), 4 + 3L for 8 (x, y,
  amplitudes, xlocs, ylocs, sigma_x, sigma_y), 5 + 3L or 6 + 3L for 10 (then
  the coupling and the optional dipole moment).

``pchip_interp`` tables (``rust/src/lindblad/plan.rs:28-115``,
``centrex_tlf/lindblad/ir.py:361-397``): the reference pairs every strictly
increasing base tuple (a grid) with every other base tuple of the same length
(values), and a call whose grid equals a table's grid (each point within
1e-15) evaluates that table -- the table's values, in Horner form -- whatever
its own values argument is.  The match is made here against the same tables;
a matched call becomes helper 17 over the table's grid and values, the others
stay helper 16 (the Hermite form of ``eval.rs:290-351``).
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

from .ir import CompiledExpression, evaluate_expression
from .ir import InstructionOp as Op
from .postfix import BINARY, COMPARE, parse, program

__all__ = ["PCHIP_TABLE_HELPER", "TUPLE_HELPERS", "extract_pchip_tables", "lower_tuples"]

TUPLE_HELPERS = frozenset({8, 10, 14, 16})
# Device-only helper id: pchip_interp through a precomputed table
# (plan.rs:33-115), arguments x, the table's grid, the table's values.
PCHIP_TABLE_HELPER = 17

_SEQUENCE_USED = "sequence value used where scalar was required"


def extract_pchip_tables(base_values: Sequence[Any]) -> list[tuple[tuple[float, ...], tuple[float, ...]]]:
    """(grid, values) pairs in the reference's order (``ir.py:361-397``)."""
    tables = []
    seen: set[tuple[int, int]] = set()
    for ci, cval in enumerate(base_values):
        if not isinstance(cval, tuple) or len(cval) < 2:
            continue
        grid = tuple(float(complex(v).real) for v in cval)
        if not all(grid[i] < grid[i + 1] for i in range(len(grid) - 1)):
            continue
        for vi, vval in enumerate(base_values):
            if vi == ci or not isinstance(vval, tuple) or len(vval) != len(cval) or (ci, vi) in seen:
                continue
            seen.add((ci, vi))
            tables.append((grid, tuple(float(complex(v).real) for v in vval)))
    return tables


def _is_const(node: tuple) -> bool:
    return node[0] == "const"


def _reads_input(node: tuple) -> bool:
    """Whether a tree reads a slot or the time."""
    kind = node[0]
    if kind in ("slot", "time"):
        return True
    if kind == "un":
        return _reads_input(node[2])
    if kind == "bin":
        return _reads_input(node[2]) or _reads_input(node[3])
    if kind == "call":
        return any(_reads_input(a) for a in node[3])
    return False


def _constant_value(node: tuple) -> float | None:
    """The real part of a tree that reads no slot and no time, or None."""
    if _is_const(node):
        return node[1]
    if _reads_input(node):
        return None
    return complex(evaluate_expression(CompiledExpression(tuple(program(node))), [], 0.0)).real


def _real_const(node: tuple, what: str) -> float:
    """The real value of a constant element (``as_real``, eval.rs:95-101)."""
    if abs(node[2]) > 1e-12:
        raise TypeError(f"complex value used where real scalar was required ({what})")
    return node[1]


class _Lowering:
    def __init__(self, base_values: Sequence[Any], n_slots: int):
        self.base = list(base_values)
        self.tables = extract_pchip_tables(self.base)
        self.next_slot = n_slots
        self.tuple_slots: dict[int, list[int]] = {}  # compound slot -> element slots

    # -- values: ("s", tree) or ("t", [trees]) --------------------------------
    def value(self, node: tuple) -> tuple[str, Any]:
        kind = node[0]
        if kind in ("const", "time", "zero"):
            return "s", node
        if kind == "slot":
            s = node[1]
            if s < len(self.base) and isinstance(self.base[s], tuple):
                return "t", [("const", complex(v).real, complex(v).imag) for v in self.base[s]]
            if s in self.tuple_slots:
                return "t", [("slot", e) for e in self.tuple_slots[s]]
            return "s", node
        if kind == "un":
            k, v = self.value(node[2])
            if k == "t":
                return "t", [("un", node[1], e) for e in v]
            return "s", ("un", node[1], v)
        if kind == "bin":
            op = node[1]
            (ka, a), (kb, b) = self.value(node[2]), self.value(node[3])
            if op in COMPARE:
                if ka == "t" or kb == "t":
                    raise TypeError(_SEQUENCE_USED)
                return "s", ("bin", op, a, b)
            if op not in BINARY:
                raise NotImplementedError(f"binary op {op}")
            if ka == "s" and kb == "s":
                return "s", ("bin", op, a, b)
            if ka == "t" and kb == "t":
                if len(a) != len(b):
                    raise ValueError("tuple lengths do not match for elementwise operation")
                return "t", [("bin", op, x, y) for x, y in zip(a, b)]
            if ka == "t":
                return "t", [("bin", op, x, b) for x in a]
            return "t", [("bin", op, a, y) for y in b]
        if kind == "call":
            op, function, args = node[1], node[2], node[3]
            vals = [self.value(a) for a in args]
            if op == Op.HELPER_FUNC and function in TUPLE_HELPERS:
                return "s", self.helper(function, vals)
            if any(k == "t" for k, _ in vals):
                raise TypeError(_SEQUENCE_USED)
            return "s", ("call", op, function, tuple(v for _, v in vals))
        raise ValueError(kind)

    # -- tuple helpers --------------------------------------------------------
    @staticmethod
    def _scalar(val: tuple[str, Any]) -> tuple:
        if val[0] != "s":
            raise TypeError(_SEQUENCE_USED)
        return val[1]

    @staticmethod
    def _seq(val: tuple[str, Any]) -> list[tuple]:
        if val[0] != "t":
            raise TypeError("helper expected tuple argument")
        return val[1]

    def helper(self, function: int, vals: list[tuple[str, Any]]) -> tuple:
        call = lambda fid, args: ("call", int(Op.HELPER_FUNC), fid, tuple(args))  # noqa: E731
        if function in (8, 10):
            if len(vals) < (7 if function == 8 else 8):
                raise ValueError(f"helper {function} takes more arguments")
            x, y = self._scalar(vals[0]), self._scalar(vals[1])
            amps, xl, yl = self._seq(vals[2]), self._seq(vals[3]), self._seq(vals[4])
            if not len(amps) == len(xl) == len(yl):
                raise ValueError("multipass helper tuple lengths do not match")
            for seq in (amps, xl, yl):
                for e in seq:
                    if _is_const(e):
                        _real_const(e, "multipass helper")
            rest = [self._scalar(v) for v in vals[5:]]
            return call(function, [x, y, *amps, *xl, *yl, *rest])
        x = self._scalar(vals[0])
        grid, values = self._seq(vals[1]), self._seq(vals[2])
        name = "linear_interp" if function == 14 else "pchip_interp"
        if len(grid) != len(values):
            raise ValueError(f"{name} grid and values must have the same length")
        if not grid:
            raise ValueError(f"{name} grid must contain at least one point")
        if function == 14:
            for e in grid:
                if _is_const(e):
                    _real_const(e, name)
        points = [_constant_value(e) for e in grid]  # None where the point is not constant
        if len(grid) > 1 and all(v is not None for v in points):
            if any(b <= a for a, b in zip(points, points[1:])):
                raise ValueError(f"{name} grid must be strictly increasing")
        if function == 16 and self.tables:
            same_length = [t for t in self.tables if len(t[0]) == len(grid)]
            if same_length and any(v is None for v in points):
                raise NotImplementedError(
                    "pchip_interp with a grid that depends on time or batch parameters, where the "
                    "problem's tuples define interpolation tables, is not lowered for the GPU")
            for tgrid, tvalues in same_length:
                if all(abs(a - b) < 1e-15 for a, b in zip(tgrid, points)):
                    return call(PCHIP_TABLE_HELPER, [x, *[("const", v, 0.0) for v in tgrid],
                                                     *[("const", v, 0.0) for v in tvalues]])
        return call(function, [x, *grid, *values])

    # -- programs -------------------------------------------------------------
    def compound(self, slot: int, expression: CompiledExpression) -> list[tuple[int, list[tuple]]]:
        if not self._touches_tuples(expression):
            return [(slot, list(expression.instructions))]
        kind, v = self.value(parse(expression.instructions))
        if kind == "s":
            return [(slot, program(v))]
        elems = list(range(self.next_slot, self.next_slot + len(v)))
        self.next_slot += len(v)
        self.tuple_slots[slot] = elems
        return [(e, program(tree)) for e, tree in zip(elems, v)]

    def coefficient(self, expression: CompiledExpression) -> list[tuple]:
        if not self._touches_tuples(expression):
            return list(expression.instructions)
        kind, v = self.value(parse(expression.instructions))
        if kind != "s":
            raise TypeError("a Hamiltonian coefficient evaluates to a sequence")
        return program(v)

    def _touches_tuples(self, expression: CompiledExpression) -> bool:
        for op, index, _argc, function, _re, _im in expression.instructions:
            if op == Op.SLOT and (index in self.tuple_slots or
                                  (index < len(self.base) and isinstance(self.base[index], tuple))):
                return True
            if op == Op.HELPER_FUNC and function in TUPLE_HELPERS:
                return True
        return False


def lower_tuples(base_values: Sequence[Any], n_slots: int,
                 compounds: Sequence[tuple[int, CompiledExpression]],
                 coefficients: Sequence[CompiledExpression]):
    """Scalar programs of a parameter graph and its coefficients.

    Returns ``(n_slots, compounds, coefficients, base)``: the slot count with
    the element slots of tuple-valued compounds, ``(slot, rows)`` for every
    compound in evaluation order (element compounds in place of a tuple
    compound), the coefficient rows, and the base values with each tuple
    replaced by 0 (a tuple slot is never read after lowering).  Programs that
    touch no tuple are returned row for row.
    """
    low = _Lowering(base_values, n_slots)
    out_compounds: list[tuple[int, list[tuple]]] = []
    for slot, expression in compounds:
        out_compounds += low.compound(slot, expression)
    out_coefs = [low.coefficient(e) for e in coefficients]
    base = [0j if isinstance(v, tuple) else v for v in base_values]
    return low.next_slot, out_compounds, out_coefs, base
