"""Parameter derivatives of the coefficient programs, built on the host.

The discrete adjoint needs d c_k(t) / d p_m for every Hamiltonian coefficient
c_k (``ir.py`` bytecode, opcodes of ``rust/src/lindblad/eval.rs:11-60``) and
every differentiated base parameter p_m (a real scalar slot such as a
detuning, Rabi rate, angle or width).  The derivatives are produced here by
symbolic differentiation of the postfix programs and emitted as ordinary
bytecode, so the device expression machine evaluates them unchanged:

* each compound slot that depends on p_m gets a derivative compound slot
  (chain rule through the parameter graph, evaluated after all compounds);
* K * P derivative programs are appended after the K coefficient programs,
  program k * P + m being d c_k / d p_m.

With ``with_time`` the time t is differentiated as one more column: the
programs are then K * (P + 1), program k * (P + 1) + P being d c_k / d t.  The
adjoint of the adaptive solve needs it because the accepted step sizes, and
so the stage times, depend on the parameters through the step-size
controller.

Rules (p real): linear ops and conj commute with d/dp; products, quotients,
powers with a parameter-free exponent, exp/sin/cos/tan by the usual
formulas; |z|' = Re(conj(z) z') / |z|; Re/Im of the derivative; comparisons
and the piecewise-constant helpers (square wave, on/off switches, alternating
sign: ids 4, 7, 12, 13) have zero derivative almost everywhere.  A power
whose exponent depends on p, or another helper whose arguments depend on p,
raises ``NotImplementedError``.
"""

from __future__ import annotations

from typing import Any, Sequence

import numpy as np

from .ir import BuiltinFunctionId as B
from .ir import InstructionOp as Op
from .postfix import COMPARE as _COMPARE
from .postfix import parse, program

__all__ = ["differentiate_flat_plan"]

_PIECEWISE_CONSTANT_HELPERS = {4, 7, 12, 13}

ZERO = ("zero",)
ONE = ("const", 1.0, 0.0)


# ---- trees -----------------------------------------------------------------

def _is_zero(n: tuple) -> bool:
    return n[0] == "zero"


def _add(a: tuple, b: tuple) -> tuple:
    if _is_zero(a):
        return b
    if _is_zero(b):
        return a
    return ("bin", Op.ADD, a, b)


def _sub(a: tuple, b: tuple) -> tuple:
    if _is_zero(b):
        return a
    if _is_zero(a):
        return ("un", Op.NEG, b)
    return ("bin", Op.SUB, a, b)


def _mul(a: tuple, b: tuple) -> tuple:
    if _is_zero(a) or _is_zero(b):
        return ZERO
    if a == ONE:
        return b
    if b == ONE:
        return a
    return ("bin", Op.MUL, a, b)


def _div(a: tuple, b: tuple) -> tuple:
    if _is_zero(a):
        return ZERO
    return ("bin", Op.DIV, a, b)


def _call(function: int, *args: tuple) -> tuple:
    return ("call", Op.BUILTIN_FUNC, int(function), tuple(args))


def _deriv(node: tuple, dslot: dict[int, tuple], wrt_time: bool = False) -> tuple:
    """d node / d p, with dslot[s] the derivative of slot s (missing = 0);
    d node / d t when ``wrt_time``."""
    kind = node[0]
    if kind == "time":
        return ONE if wrt_time else ZERO
    if kind in ("zero", "const"):
        return ZERO
    if kind == "slot":
        return dslot.get(node[1], ZERO)
    if kind == "un":
        d = _deriv(node[2], dslot, wrt_time)
        return ZERO if _is_zero(d) else ("un", node[1], d)
    if kind == "bin":
        op, a, b = node[1], node[2], node[3]
        if op in _COMPARE:
            return ZERO
        da, db = _deriv(a, dslot, wrt_time), _deriv(b, dslot, wrt_time)
        if op == Op.ADD:
            return _add(da, db)
        if op == Op.SUB:
            return _sub(da, db)
        if op == Op.MUL:
            return _add(_mul(da, b), _mul(a, db))
        if op == Op.DIV:
            return _sub(_div(da, b), _div(_mul(a, db), _mul(b, b)))
        if op == Op.POW:
            if not _is_zero(db):
                raise NotImplementedError("derivative of a power whose exponent depends on the parameter")
            if _is_zero(da):
                return ZERO
            # b * a^(b - 1) * a'
            bm1 = ("const", b[1] - 1.0, b[2]) if b[0] == "const" else ("bin", Op.SUB, b, ONE)
            return _mul(_mul(b, ("bin", Op.POW, a, bm1)), da)
        raise NotImplementedError(f"binary op {op}")
    if kind == "call":
        op, function, args = node[1], node[2], node[3]
        dargs = [_deriv(a, dslot, wrt_time) for a in args]
        if all(_is_zero(d) for d in dargs):
            return ZERO
        if op == Op.HELPER_FUNC:
            if function in _PIECEWISE_CONSTANT_HELPERS:
                return ZERO
            raise NotImplementedError(f"parameter derivative of helper function {function} is not supported")
        a, da = args[0], dargs[0]
        if function == B.EXP:
            return _mul(_call(B.EXP, a), da)
        if function == B.SIN:
            return _mul(_call(B.COS, a), da)
        if function == B.COS:
            return _mul(("un", Op.NEG, _call(B.SIN, a)), da)
        if function == B.TAN:
            t = _call(B.TAN, a)
            return _mul(("bin", Op.ADD, ONE, ("bin", Op.MUL, t, t)), da)
        if function == B.ABS:
            return _div(_call(B.REAL, _mul(("un", Op.CONJ, a), da)), _call(B.ABS, a))
        if function == B.REAL:
            return _call(B.REAL, da)
        if function == B.IMAG:
            return _call(B.IMAG, da)
        raise NotImplementedError(f"builtin function {function}")
    raise ValueError(kind)


def differentiate_flat_plan(flat: dict[str, Any], param_slots: Sequence[int],
                            with_time: bool = False) -> dict[str, Any]:
    """A copy of ``flat`` with derivative compounds and K*P derivative programs.

    ``param_slots`` are base-slot indices (the differentiated parameters).  The
    returned plan has ``n_diff_params = P`` and ``diff_param_slots``; its
    expression list is compounds (original, then derivative compounds),
    the K coefficients, then the derivative programs (k * P' + m, P' = P, or
    P + 1 with ``with_time``: column P is d c_k / d t and ``diff_time = 1``).
    """
    n_base = int(flat["n_base"])
    params = [int(s) for s in param_slots]
    for s in params:
        if not 0 <= s < n_base:
            raise ValueError(f"parameter slot {s} is not a base parameter slot")
    ptrs = np.asarray(flat["expr_ptrs"])
    rows = list(zip(flat["instr_op"], flat["instr_index"], flat["instr_argc"], flat["instr_function"],
                    flat["instr_re"], flat["instr_im"]))
    n_comp, K = int(flat["n_compounds"]), int(flat["n_coefficients"])
    trees = [parse(rows[ptrs[e]:ptrs[e + 1]]) for e in range(n_comp + K)]
    comp_slots = [int(s) for s in flat["compound_slots"]]
    next_slot = int(flat["n_slots"])
    deriv_compounds: list[tuple[int, list[tuple]]] = []
    dslots: list[dict[int, tuple]] = []
    columns = [(p, False) for p in params] + ([(None, True)] if with_time else [])
    for p, wrt_time in columns:
        dslot: dict[int, tuple] = {} if wrt_time else {p: ONE}
        for c in range(n_comp):
            d = _deriv(trees[c], dslot, wrt_time)
            if _is_zero(d):
                continue
            if d[0] == "const" or d[0] == "slot":
                dslot[comp_slots[c]] = d
                continue
            dslot[comp_slots[c]] = ("slot", next_slot)
            deriv_compounds.append((next_slot, program(d)))
            next_slot += 1
        dslots.append(dslot)
    deriv_coefs = [program(_deriv(trees[n_comp + k], dslots[m], columns[m][1]))
                   for k in range(K) for m in range(len(columns))]

    programs = [rows[ptrs[e]:ptrs[e + 1]] for e in range(n_comp)]
    programs += [p for _, p in deriv_compounds]
    programs += [rows[ptrs[n_comp + k]:ptrs[n_comp + k + 1]] for k in range(K)]
    programs += deriv_coefs
    out = dict(flat)
    flat_rows = [r for p in programs for r in p]
    ins = np.asarray(flat_rows, dtype=np.float64).reshape(-1, 6)
    new_ptrs = np.zeros(len(programs) + 1, dtype=np.int32)
    np.cumsum([len(p) for p in programs], out=new_ptrs[1:])
    out.update({
        "n_slots": next_slot,
        "n_compounds": n_comp + len(deriv_compounds),
        "compound_slots": np.asarray(comp_slots + [s for s, _ in deriv_compounds], dtype=np.int32),
        "expr_ptrs": new_ptrs,
        "instr_op": np.ascontiguousarray(ins[:, 0].astype(np.int32)),
        "instr_index": np.ascontiguousarray(ins[:, 1].astype(np.int32)),
        "instr_argc": np.ascontiguousarray(ins[:, 2].astype(np.int32)),
        "instr_function": np.ascontiguousarray(ins[:, 3].astype(np.int32)),
        "instr_re": np.ascontiguousarray(ins[:, 4]),
        "instr_im": np.ascontiguousarray(ins[:, 5]),
        "n_diff_params": len(params),
        "diff_param_slots": np.asarray(params, dtype=np.int32),
        "diff_time": int(bool(with_time)),
    })
    return out
