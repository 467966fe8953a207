"""Expression trees of the postfix bytecode (``ir.py`` rows, opcodes of
``rust/src/lindblad/eval.rs:11-60``), for the host passes that rewrite
programs (parameter derivatives, tuple lowering).

A tree is a nested tuple: ``("const", re, im)``, ``("slot", index)``,
``("time",)``, ``("un", op, a)`` (NEG, CONJ), ``("bin", op, a, b)``
(arithmetic and comparisons) or ``("call", op, function, args)`` (builtins and
helpers).  :func:`emit` writes a tree back in the order :func:`parse` read it,
so an unchanged tree reproduces its program row for row.
"""

from __future__ import annotations

from collections.abc import Sequence

from .ir import InstructionOp as Op

__all__ = ["BINARY", "COMPARE", "emit", "parse", "program"]

BINARY = {Op.ADD, Op.SUB, Op.MUL, Op.DIV, Op.POW}
COMPARE = set(range(14, 20))


def parse(rows: Sequence[tuple]) -> tuple:
    """The tree of one program (rows of op, index, argc, function, re, im)."""
    stack: list[tuple] = []
    for op, index, argc, function, re, im in rows:
        op, index, argc, function = int(op), int(index), int(argc), int(function)
        if op == Op.CONST:
            stack.append(("const", float(re), float(im)))
        elif op == Op.SLOT:
            stack.append(("slot", index))
        elif op == Op.TIME:
            stack.append(("time",))
        elif op in BINARY or op in COMPARE:
            b, a = stack.pop(), stack.pop()
            stack.append(("bin", op, a, b))
        elif op in (Op.NEG, Op.CONJ):
            stack.append(("un", op, stack.pop()))
        elif op in (Op.BUILTIN_FUNC, Op.HELPER_FUNC):
            args = tuple(stack[len(stack) - argc:])
            del stack[len(stack) - argc:]
            stack.append(("call", op, function, args))
        else:
            raise NotImplementedError(f"instruction op {op} has no tree form")
    if len(stack) != 1:
        raise ValueError("malformed expression")
    return stack[0]


def emit(node: tuple, out: list[tuple]) -> None:
    """Append the rows of ``node`` to ``out``."""
    kind = node[0]
    if kind == "zero":
        out.append((int(Op.CONST), 0, 0, 0, 0.0, 0.0))
    elif kind == "const":
        out.append((int(Op.CONST), 0, 0, 0, node[1], node[2]))
    elif kind == "slot":
        out.append((int(Op.SLOT), node[1], 0, 0, 0.0, 0.0))
    elif kind == "time":
        out.append((int(Op.TIME), 0, 0, 0, 0.0, 0.0))
    elif kind == "bin":
        emit(node[2], out)
        emit(node[3], out)
        out.append((int(node[1]), 0, 0, 0, 0.0, 0.0))
    elif kind == "un":
        emit(node[2], out)
        out.append((int(node[1]), 0, 0, 0, 0.0, 0.0))
    elif kind == "call":
        for a in node[3]:
            emit(a, out)
        out.append((int(node[1]), 0, len(node[3]), int(node[2]), 0.0, 0.0))
    else:
        raise ValueError(kind)


def program(node: tuple) -> list[tuple]:
    """The rows of ``node``."""
    out: list[tuple] = []
    emit(node, out)
    return out
