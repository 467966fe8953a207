"""Minimal OBE system container and loaders for serialised systems.

The reference builds ``OBESystem`` objects from its molecular builders
(``centrex_tlf/lindblad/utils_setup.py:50-156``).  The solver reads only
``H_symbolic`` (a sympy matrix over parameter symbols) and ``C_array``
(single-jump collapse operators), so that is what this container holds, plus
state labels.  :func:`load_obe_system` reads the JSON written by
``tests/golden/make_tlf_systems.py``.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import sympy as smp

__all__ = ["OBESystem", "load_obe_system", "two_level_system"]


@dataclass
class OBESystem:
    H_symbolic: smp.Matrix
    C_array: np.ndarray
    QN: list[dict[str, Any]] = field(default_factory=list)
    description: str = ""
    gamma: float | None = None

    @property
    def n_states(self) -> int:
        return int(self.H_symbolic.shape[0])

    def indices(self, **selector: Any) -> list[int]:
        """State indices whose labels match every ``key=value`` in ``selector``."""
        out = []
        for idx, label in enumerate(self.QN):
            if all(label.get(k) == v for k, v in selector.items()):
                out.append(idx)
        return out


def load_obe_system(path: str | Path) -> OBESystem:
    data = json.loads(Path(path).read_text())
    n = int(data["n"])
    namespace = {name: getattr(smp, name) for name in dir(smp) if not name.startswith("_")}
    H = smp.zeros(n, n)
    for i, j, expr in data["hamiltonian_entries"]:
        H[int(i), int(j)] = eval(expr, {"__builtins__": {}}, namespace)  # sympy srepr text
    C = np.zeros((int(data["n_collapse"]), n, n), dtype=np.complex128)
    for k, target, source, re, im in data["collapse_jumps"]:
        C[int(k), int(target), int(source)] = complex(re, im)
    return OBESystem(H_symbolic=H, C_array=C, QN=list(data.get("states", [])),
                     description=str(data.get("description", "")), gamma=data.get("gamma"))


def two_level_system(decay: float = 0.3, omega_name: str = "Ω", delta_name: str = "δ") -> OBESystem:
    """H = [[0, Ω/2], [Ω*/2, -δ]] with one decay channel |1> -> |0> of rate ``decay``.

    The fixture of ``tests/lindblad/test_rust_backend.py:35-58`` and of the
    reference's golden benchmarks (``benchmarks/bench_expanded_sparse_packed_rhs_kernels.py:100-125``).
    """
    omega, delta = smp.symbols(f"{omega_name} {delta_name}", real=True)
    H = smp.Matrix([[0, omega / 2], [smp.conjugate(omega) / 2, -delta]])
    C = np.zeros((1, 2, 2), dtype=np.complex128)
    C[0, 0, 1] = np.sqrt(decay)
    return OBESystem(H_symbolic=H, C_array=C, QN=[{"index": 0}, {"index": 1}],
                     description="two-level system")
