"""Host lowering of an OBE system into the flat tables the GPU solver consumes.

Three lowerings, each following the reference's semantics so that the GPU path
integrates exactly the same right-hand side:

* :func:`compile_expression` turns a sympy expression into stack bytecode with
  the reference's opcode, builtin and helper ids
  (``centrex_tlf/lindblad/ir.py:28-67, 210-326``; ``rust/src/lindblad/eval.rs:11-153``).
* :func:`lower_hamiltonian_decomposed` splits the upper triangle of the
  symbolic Hamiltonian into a static matrix plus ``K`` symbolic coefficients
  times numeric basis matrices (``ir.py:677-879``).
* :func:`lower_expanded_sparse_rhs` emits the CSR "program" over the
  ``n(n+1)/2`` upper-triangle outputs: every term is
  ``factor * (c_k or conj(c_k) or 1) * (rho_in or conj(rho_in))``; the
  commutator ``-i[H, rho]`` and the single-jump dissipator are folded into the
  same terms (``ir.py:447-613``).  Keys are accumulated and sorted exactly as
  the reference does so the factors agree bit for bit.

:func:`flatten_plan` packs everything into numpy arrays for the C ABI
(``include/centrex_lindblad.h``, ``lb_plan_desc``).
"""

from __future__ import annotations

import cmath
from collections import defaultdict
from dataclasses import dataclass
from enum import IntEnum
from typing import Any

import numpy as np
import sympy as smp
from sympy.core.relational import Relational

from .helper_functions import HELPER_FUNCTION_IDS, HELPER_FUNCTION_NAMES, HELPER_FUNCTIONS
from .state_layout import upper_offsets

__all__ = [
    "BuiltinFunctionId",
    "CompiledExpression",
    "InstructionOp",
    "compile_expression",
    "evaluate_expression",
    "evaluate_parameter_graph",
    "flatten_plan",
    "lower_expanded_sparse_rhs",
    "lower_hamiltonian_decomposed",
    "lower_parameter_graph",
]


class InstructionOp(IntEnum):
    CONST = 1
    SLOT = 2
    TEMP = 3
    TIME = 4
    ADD = 5
    SUB = 6
    MUL = 7
    DIV = 8
    POW = 9
    NEG = 10
    CONJ = 11
    BUILTIN_FUNC = 12
    HELPER_FUNC = 13
    GT = 14
    GE = 15
    LT = 16
    LE = 17
    EQ = 18
    NE = 19


class BuiltinFunctionId(IntEnum):
    SIN = 1
    COS = 2
    TAN = 3
    EXP = 4
    ABS = 5
    REAL = 6
    IMAG = 7


BUILTIN_FUNCTION_IDS = {
    "sin": BuiltinFunctionId.SIN,
    "cos": BuiltinFunctionId.COS,
    "tan": BuiltinFunctionId.TAN,
    "exp": BuiltinFunctionId.EXP,
    "Abs": BuiltinFunctionId.ABS,
    "re": BuiltinFunctionId.REAL,
    "im": BuiltinFunctionId.IMAG,
}

_RELATIONAL_OPS = {
    "StrictGreaterThan": InstructionOp.GT,
    "GreaterThan": InstructionOp.GE,
    "StrictLessThan": InstructionOp.LT,
    "LessThan": InstructionOp.LE,
    "Equality": InstructionOp.EQ,
    "Unequality": InstructionOp.NE,
}


@dataclass(frozen=True)
class CompiledExpression:
    """Postfix program: list of ``(op, index, argc, function, re, im)`` rows."""

    instructions: tuple[tuple[int, int, int, int, float, float], ...]
    scalar_only: bool = True
    output_is_tuple: bool = False

    @property
    def uses_time(self) -> bool:
        return any(ins[0] == InstructionOp.TIME for ins in self.instructions)

    def slot_reads(self) -> set[int]:
        return {ins[1] for ins in self.instructions if ins[0] == InstructionOp.SLOT}


def _const(value: complex) -> tuple[int, int, int, int, float, float]:
    return (int(InstructionOp.CONST), 0, 0, 0, float(value.real), float(value.imag))


def compile_expression(expr: smp.Basic, slot_index_by_name: dict[str, int],
                       tuple_value_names: set[str] | None = None) -> CompiledExpression:
    """Compile a sympy expression to postfix bytecode (``ir.py:210-326``).

    Add and Mul chains are emitted left to right over ``node.args`` with one
    binary op per extra argument, so the evaluation order (and hence the
    rounding) is the reference's.
    """
    tuple_value_names = tuple_value_names or set()
    out: list[tuple[int, int, int, int, float, float]] = []

    def op(code: InstructionOp, index: int = 0, argc: int = 0, function: int = 0) -> None:
        out.append((int(code), int(index), int(argc), int(function), 0.0, 0.0))

    def emit(node: smp.Basic) -> tuple[bool, bool]:
        if node == smp.I:
            out.append(_const(1j))
            return True, False
        if isinstance(node, smp.Symbol):
            name = str(node)
            if name == "t":
                op(InstructionOp.TIME)
                return True, False
            if name in slot_index_by_name:
                op(InstructionOp.SLOT, slot_index_by_name[name])
                is_tuple = name in tuple_value_names
                return not is_tuple, is_tuple
            raise KeyError(f"symbol {name} not found in slot namespace")
        if isinstance(node, smp.Number) or node.is_number:
            out.append(_const(complex(node.evalf())))
            return True, False
        if isinstance(node, Relational):
            ls, lt = emit(node.lhs)
            rs, rt = emit(node.rhs)
            try:
                code = _RELATIONAL_OPS[type(node).__name__]
            except KeyError as exc:
                raise NotImplementedError(f"unsupported relational {type(node).__name__}") from exc
            op(code)
            return ls and rs and not (lt or rt), False
        if isinstance(node, smp.Add | smp.Mul):
            code = InstructionOp.ADD if isinstance(node, smp.Add) else InstructionOp.MUL
            args = list(node.args)
            scalar, tup = emit(args[0])
            for arg in args[1:]:
                s2, t2 = emit(arg)
                op(code)
                scalar, tup = scalar and s2, tup or t2
            return scalar and not tup, tup
        if isinstance(node, smp.Pow):
            ls, lt = emit(node.args[0])
            rs, rt = emit(node.args[1])
            op(InstructionOp.POW)
            tup = lt or rt
            return ls and rs and not tup, tup
        if node.func == smp.conjugate:
            s, t = emit(node.args[0])
            op(InstructionOp.CONJ)
            return s and not t, t
        if node.is_Function:
            name = getattr(node.func, "__name__", str(node.func))
            scalar, saw_tuple = True, False
            for arg in node.args:
                s, t = emit(arg)
                scalar, saw_tuple = scalar and s, saw_tuple or t
            if name in BUILTIN_FUNCTION_IDS:
                op(InstructionOp.BUILTIN_FUNC, argc=len(node.args), function=BUILTIN_FUNCTION_IDS[name])
                return scalar and not saw_tuple, False
            if name in HELPER_FUNCTION_IDS:
                op(InstructionOp.HELPER_FUNC, argc=len(node.args), function=HELPER_FUNCTION_IDS[name])
                return scalar and not saw_tuple, False
            raise NotImplementedError(f"unsupported function {name}")
        raise NotImplementedError(f"unsupported sympy node {type(node).__name__}")

    scalar_only, is_tuple = emit(smp.sympify(expr))
    return CompiledExpression(tuple(out), scalar_only, is_tuple)


def lower_parameter_graph(parameters: Any) -> dict[str, Any]:
    """Slots, base values and compound bytecode (``ir.py:329-358``)."""
    slot_index = parameters.slot_index_by_name
    tuple_names = {k for k, v in parameters.base_parameters.items() if isinstance(v, tuple)}
    compounds = []
    for name, expr in parameters.parsed_compound_parameters.items():
        compiled = compile_expression(expr, slot_index, tuple_names)
        compounds.append((slot_index[name], compiled))
        if compiled.output_is_tuple:
            tuple_names.add(str(name))
    return {
        "slot_names": list(parameters.slot_names),
        "base_values": list(parameters.base_parameters.values()),
        "compounds": compounds,
        "tuple_value_names": tuple_names,
    }


def lower_hamiltonian_decomposed(hamiltonian: smp.Matrix, slot_index_by_name: dict[str, int],
                                 tuple_value_names: set[str] | None = None) -> dict[str, Any]:
    """Static matrix + symbolic coefficients x numeric upper-triangle bases (``ir.py:677-879``).

    Every upper entry is simplified, expanded and split term by term into a
    numeric factor and a symbolic remainder; remainders are de-duplicated by
    ``srepr`` in first-seen order, which fixes the coefficient numbering.
    """
    n = int(hamiltonian.rows)
    static = np.zeros((n, n), dtype=np.complex128)
    terms: dict[str, tuple[smp.Expr, np.ndarray]] = {}

    def split(term: smp.Expr) -> tuple[complex, smp.Expr]:
        if term == 0:
            return 0j, smp.Integer(1)
        if term.is_number:
            return complex(term.evalf()), smp.Integer(1)
        numeric = smp.Integer(1)
        symbolic = []
        for factor in smp.Mul.make_args(term):
            if factor.is_number:
                numeric *= factor
            else:
                symbolic.append(factor)
        sym = smp.Mul(*symbolic) if symbolic else smp.Integer(1)
        return complex(numeric.evalf()), smp.simplify(smp.factor_terms(sym))

    for i in range(n):
        for j in range(i, n):
            entry = hamiltonian[i, j]
            if entry == 0:
                continue
            expr = smp.expand(smp.factor_terms(smp.simplify(entry)))
            for term in smp.Add.make_args(expr):
                numeric, symbolic = split(term)
                if abs(numeric) == 0.0:
                    continue
                if symbolic == 1:
                    static[i, j] += numeric
                    if i != j:
                        static[j, i] += np.conjugate(numeric)
                    continue
                key = smp.srepr(symbolic)
                if key not in terms:
                    terms[key] = (symbolic, np.zeros((n, n), dtype=np.complex128))
                terms[key][1][i, j] += numeric

    coefficients = []
    for symbolic, basis in terms.values():
        compiled = compile_expression(symbolic, slot_index_by_name, tuple_value_names)
        ii, jj = np.nonzero(np.triu(basis))
        coefficients.append({
            "expression": compiled,
            "symbolic": symbolic,
            "basis_terms": [(int(a), int(b), complex(basis[a, b])) for a, b in zip(ii, jj)],
        })
    return {"kind": "decomposed", "n": n, "static_matrix": static, "coefficients": coefficients}


def lower_hamiltonian_entrywise(hamiltonian: smp.Matrix, slot_index_by_name: dict[str, int],
                                tuple_value_names: set[str] | None = None) -> dict[str, Any]:
    """Entrywise Hamiltonian plan (``ir.py:616-674``): every nonzero upper entry
    H_ij is evaluated from its own simplified expression (numbers go to the
    static matrix).  The entries become coefficients of the expanded sparse
    program with one unit basis term each, so the GPU runs them like a
    decomposed plan; equal entry expressions share one coefficient (the same
    value either way).  The entries are not split into numeric and symbolic
    factors as the decomposed plan does, so the values are rounded as the
    entry expressions themselves are.
    """
    n = int(hamiltonian.rows)
    static = np.zeros((n, n), dtype=np.complex128)
    entries: dict[str, tuple[smp.Expr, np.ndarray]] = {}
    for i in range(n):
        for j in range(i, n):
            expr = smp.simplify(hamiltonian[i, j])
            if expr == 0:
                continue
            if expr.is_number:
                static[i, j] += complex(expr.evalf())
                if i != j:
                    static[j, i] += np.conjugate(complex(expr.evalf()))
                continue
            key = smp.srepr(expr)
            if key not in entries:
                entries[key] = (expr, np.zeros((n, n), dtype=np.complex128))
            entries[key][1][i, j] = 1.0
    coefficients = []
    for expr, basis in entries.values():
        ii, jj = np.nonzero(np.triu(basis))
        coefficients.append({
            "expression": compile_expression(expr, slot_index_by_name, tuple_value_names),
            "symbolic": expr,
            "basis_terms": [(int(a), int(b), 1.0 + 0.0j) for a, b in zip(ii, jj)],
        })
    return {"kind": "entrywise", "n": n, "static_matrix": static, "coefficients": coefficients}


def lower_expanded_sparse_rhs(hamiltonian_plan: dict[str, Any], source_decay_rates: np.ndarray,
                              incoming_transfers_by_target: list[list[dict[str, Any]]]) -> dict[str, Any]:
    """CSR program of the upper-triangle Lindblad RHS (``ir.py:447-613``).

    Output ``(i, j)`` collects ``-i H_ik rho_kj`` (row terms), ``+i rho_ik H_kj``
    (column terms), ``-(g_i + g_j)/2 rho_ij`` off the diagonal and
    ``-g_i rho_ii + sum_s rate_{s->i} rho_ss`` on it.  Terms with the same
    ``(input, coefficient, coefficient_conj, input_conj)`` key are summed, keys
    are emitted in sorted order and ``|factor| <= 1e-15`` is dropped.
    """
    n = int(hamiltonian_plan["n"])
    offsets = upper_offsets(n)
    upper_len = n * (n + 1) // 2

    def uidx(i: int, j: int) -> int:
        return int(offsets[i]) + (j - i)

    h_rows: list[list[tuple[int, int, bool, complex]]] = [[] for _ in range(n)]
    h_cols: list[list[tuple[int, int, bool, complex]]] = [[] for _ in range(n)]

    def add_full(row: int, col: int, coeff: int, conj: bool, factor: complex) -> None:
        if factor == 0:
            return
        h_rows[row].append((col, coeff, conj, factor))
        h_cols[col].append((row, coeff, conj, factor))

    def add_upper(row: int, col: int, coeff: int, factor: complex) -> None:
        add_full(row, col, coeff, False, factor)
        if row != col:
            add_full(col, row, coeff, coeff >= 0, complex(np.conjugate(factor)))

    static = np.asarray(hamiltonian_plan["static_matrix"], dtype=np.complex128)
    for i in range(n):
        for j in range(i, n):
            value = complex(static[i, j])
            if value != 0:
                add_upper(i, j, -1, value)
    for k, coefficient in enumerate(hamiltonian_plan["coefficients"]):
        for i, j, value in coefficient["basis_terms"]:
            if value != 0:
                add_upper(i, j, k, complex(value))

    decay = np.asarray(source_decay_rates, dtype=np.float64)
    by_output: list[dict[tuple[int, int, bool, bool], complex]] = [
        defaultdict(complex) for _ in range(upper_len)
    ]

    def add_term(oi: int, oj: int, ii: int, ij: int, coeff: int, cconj: bool, factor: complex) -> None:
        if factor == 0:
            return
        if ii <= ij:
            key = (uidx(ii, ij), coeff, cconj, False)
        else:
            key = (uidx(ij, ii), coeff, cconj, True)
        by_output[uidx(oi, oj)][key] += factor

    for i in range(n):
        for j in range(i, n):
            for k, coeff, cconj, h in h_rows[i]:
                add_term(i, j, k, j, coeff, cconj, -1j * h)
            for k, coeff, cconj, h in h_cols[j]:
                add_term(i, j, i, k, coeff, cconj, 1j * h)
            if i == j:
                if decay[i] != 0.0:
                    add_term(i, j, i, i, -1, False, complex(-float(decay[i])))
                for transfer in incoming_transfers_by_target[i]:
                    rate = float(transfer["rate"])
                    if rate != 0.0:
                        src = int(transfer["source"])
                        add_term(i, j, src, src, -1, False, complex(rate))
            else:
                rate = 0.5 * float(decay[i] + decay[j])
                if rate != 0.0:
                    add_term(i, j, i, j, -1, False, complex(-rate))

    output_ptrs = [0]
    cols: dict[str, list] = {k: [] for k in ("input", "coeff", "cconj", "iconj", "fre", "fim")}
    for terms in by_output:
        for (inp, coeff, cconj, iconj), factor in sorted(terms.items()):
            if abs(factor) <= 1e-15:
                continue
            cols["input"].append(inp)
            cols["coeff"].append(coeff)
            cols["cconj"].append(cconj)
            cols["iconj"].append(iconj)
            cols["fre"].append(factor.real)
            cols["fim"].append(factor.imag)
        output_ptrs.append(len(cols["input"]))
    return {
        "kind": "expanded_sparse",
        "n": n,
        "output_ptrs": np.asarray(output_ptrs, dtype=np.int64),
        "input_indices": np.asarray(cols["input"], dtype=np.int64),
        "coeff_indices": np.asarray(cols["coeff"], dtype=np.int64),
        "coeff_conj": np.asarray(cols["cconj"], dtype=np.uint8),
        "input_conj": np.asarray(cols["iconj"], dtype=np.uint8),
        "factors_re": np.asarray(cols["fre"], dtype=np.float64),
        "factors_im": np.asarray(cols["fim"], dtype=np.float64),
        "diagnostics": {
            "upper_len": upper_len,
            "term_count": len(cols["input"]),
            "hamiltonian_full_term_count": sum(len(r) for r in h_rows),
            "coefficient_count": len(hamiltonian_plan["coefficients"]),
        },
    }


# ---------------------------------------------------------------------------
# Host evaluation of bytecode (used for plan checks and host-side plotting).
# ---------------------------------------------------------------------------


def _as_real(value: Any) -> float:
    if isinstance(value, tuple):
        raise TypeError("sequence value used where scalar was required")
    if abs(value.imag) > 1e-12:
        raise TypeError("complex value used where real scalar was required")
    return float(value.real)


def _elementwise(left: Any, right: Any, fn) -> Any:
    if isinstance(left, tuple) and isinstance(right, tuple):
        if len(left) != len(right):
            raise ValueError("tuple lengths do not match for elementwise operation")
        return tuple(fn(a, b) for a, b in zip(left, right))
    if isinstance(left, tuple):
        return tuple(fn(a, right) for a in left)
    if isinstance(right, tuple):
        return tuple(fn(left, b) for b in right)
    return fn(left, right)


def _powc(base: complex, exponent: complex) -> complex:
    # exp(exponent * log(base)) in polar form; 0**0 = 1 and 0**w = 0.
    if exponent == 0:
        return 1 + 0j
    r, theta = abs(base), cmath.phase(base)
    if r == 0.0:
        return 0j
    import math

    mag = r**exponent.real * math.exp(-exponent.imag * theta)
    ang = exponent.real * theta + exponent.imag * math.log(r)
    return complex(mag * math.cos(ang), mag * math.sin(ang))


def evaluate_expression(expression: CompiledExpression, slots: list[Any], t: float) -> Any:
    """Run bytecode on the host (``eval.rs:500-656`` semantics)."""
    stack: list[Any] = []
    for code, index, argc, function, re, im in expression.instructions:
        if code == InstructionOp.CONST:
            stack.append(complex(re, im))
        elif code == InstructionOp.SLOT:
            stack.append(slots[index])
        elif code == InstructionOp.TIME:
            stack.append(complex(t, 0.0))
        elif code in (InstructionOp.ADD, InstructionOp.SUB, InstructionOp.MUL, InstructionOp.DIV,
                      InstructionOp.POW):
            right, left = stack.pop(), stack.pop()
            fn = {
                InstructionOp.ADD: lambda a, b: a + b,
                InstructionOp.SUB: lambda a, b: a - b,
                InstructionOp.MUL: lambda a, b: a * b,
                InstructionOp.DIV: lambda a, b: a / b,
                InstructionOp.POW: _powc,
            }[InstructionOp(code)]
            stack.append(_elementwise(left, right, fn))
        elif code == InstructionOp.NEG:
            v = stack.pop()
            stack.append(tuple(-a for a in v) if isinstance(v, tuple) else -v)
        elif code == InstructionOp.CONJ:
            v = stack.pop()
            stack.append(tuple(a.conjugate() for a in v) if isinstance(v, tuple) else v.conjugate())
        elif code == InstructionOp.BUILTIN_FUNC:
            args = stack[len(stack) - argc:]
            del stack[len(stack) - argc:]
            v = args[0]
            bid = BuiltinFunctionId(function)
            stack.append({
                BuiltinFunctionId.SIN: lambda: cmath.sin(v),
                BuiltinFunctionId.COS: lambda: cmath.cos(v),
                BuiltinFunctionId.TAN: lambda: cmath.tan(v),
                BuiltinFunctionId.EXP: lambda: cmath.exp(v),
                BuiltinFunctionId.ABS: lambda: complex(abs(v), 0.0),
                BuiltinFunctionId.REAL: lambda: complex(v.real, 0.0),
                BuiltinFunctionId.IMAG: lambda: complex(v.imag, 0.0),
            }[bid]())
        elif code == InstructionOp.HELPER_FUNC:
            args = stack[len(stack) - argc:]
            del stack[len(stack) - argc:]
            coerced = [tuple(_as_real(a) for a in arg) if isinstance(arg, tuple) else
                       (arg if abs(arg.imag) > 1e-12 else float(arg.real)) for arg in args]
            stack.append(complex(HELPER_FUNCTIONS[HELPER_FUNCTION_NAMES[function]](*coerced)))
        elif code in (InstructionOp.GT, InstructionOp.GE, InstructionOp.LT, InstructionOp.LE,
                      InstructionOp.EQ, InstructionOp.NE):
            right, left = _as_real(stack.pop()), _as_real(stack.pop())
            res = {
                InstructionOp.GT: left > right, InstructionOp.GE: left >= right,
                InstructionOp.LT: left < right, InstructionOp.LE: left <= right,
                InstructionOp.EQ: left == right, InstructionOp.NE: left != right,
            }[InstructionOp(code)]
            stack.append(complex(1.0 if res else 0.0, 0.0))
        else:
            raise ValueError(f"unsupported instruction op {code}")
    if len(stack) != 1:
        raise ValueError("expression evaluation did not end with a single stack value")
    return stack[0]


def evaluate_parameter_graph(graph: dict[str, Any], t: float,
                             overrides: dict[int, complex] | None = None) -> list[Any]:
    slots: list[Any] = list(graph["base_values"])
    for slot, value in (overrides or {}).items():
        slots[slot] = complex(value)
    slots.extend([0j] * (len(graph["slot_names"]) - len(slots)))
    for slot, expression in graph["compounds"]:
        slots[slot] = evaluate_expression(expression, slots, t)
    return slots


# ---------------------------------------------------------------------------
# Flattening for the C ABI.
# ---------------------------------------------------------------------------


def flatten_plan(n: int, parameter_graph: dict[str, Any], hamiltonian_plan: dict[str, Any],
                 rhs_plan: dict[str, Any]) -> dict[str, np.ndarray | int]:
    """Structure-of-arrays view of a prepared plan, as ``lb_plan_desc`` expects.

    Expressions are concatenated: compounds first (in evaluation order), then
    the Hamiltonian coefficients; ``expr_ptrs`` delimits them.  Tuple-valued
    parameters are lowered to scalar programs first (``tuples.lower_tuples``):
    the device expression machine is scalar.
    """
    from .tuples import lower_tuples

    n_slots, compounds, coefs, base = lower_tuples(
        parameter_graph["base_values"], len(parameter_graph["slot_names"]), parameter_graph["compounds"],
        [c["expression"] for c in hamiltonian_plan["coefficients"]])
    exprs = [rows for _, rows in compounds] + coefs
    rows = [ins for e in exprs for ins in e]
    ptrs = np.zeros(len(exprs) + 1, dtype=np.int32)
    np.cumsum([len(e) for e in exprs], out=ptrs[1:])
    ins = np.asarray(rows, dtype=np.float64).reshape(-1, 6) if rows else np.zeros((0, 6))
    base_arr = np.asarray([complex(v) for v in base], dtype=np.complex128)
    return {
        "n_states": int(n),
        "n_slots": int(n_slots),
        "n_base": len(base),
        "base_re": np.ascontiguousarray(base_arr.real),
        "base_im": np.ascontiguousarray(base_arr.imag),
        "n_compounds": len(compounds),
        "compound_slots": np.asarray([s for s, _ in compounds], dtype=np.int32),
        "n_coefficients": len(hamiltonian_plan["coefficients"]),
        "expr_ptrs": ptrs,
        "instr_op": np.ascontiguousarray(ins[:, 0].astype(np.int32)),
        "instr_index": np.ascontiguousarray(ins[:, 1].astype(np.int32)),
        "instr_argc": np.ascontiguousarray(ins[:, 2].astype(np.int32)),
        "instr_function": np.ascontiguousarray(ins[:, 3].astype(np.int32)),
        "instr_re": np.ascontiguousarray(ins[:, 4]),
        "instr_im": np.ascontiguousarray(ins[:, 5]),
        "output_ptrs": np.ascontiguousarray(rhs_plan["output_ptrs"], dtype=np.int64),
        "input_indices": np.ascontiguousarray(rhs_plan["input_indices"], dtype=np.int64),
        "coeff_indices": np.ascontiguousarray(rhs_plan["coeff_indices"], dtype=np.int64),
        "coeff_conj": np.ascontiguousarray(rhs_plan["coeff_conj"], dtype=np.uint8),
        "input_conj": np.ascontiguousarray(rhs_plan["input_conj"], dtype=np.uint8),
        "factors_re": np.ascontiguousarray(rhs_plan["factors_re"], dtype=np.float64),
        "factors_im": np.ascontiguousarray(rhs_plan["factors_im"], dtype=np.float64),
    }
