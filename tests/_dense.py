"""Textbook dense Lindblad RHS in numpy, used as an independent check of the
lowered sparse program: drho/dt = -i[H, rho] + sum_k (C_k rho C_k^+ - {C_k^+ C_k, rho}/2).

The numeric Hamiltonian is obtained by sympy substitution + lambdify of the
parameter expressions, so it does not go through this project's bytecode.
"""

from __future__ import annotations

import numpy as np
import sympy as smp

from centrex_tlf_amd.lindblad.helper_functions import HELPER_FUNCTIONS


def slot_values(prepared, t: float, overrides: dict | None = None) -> dict[str, complex]:
    params = prepared.parameters
    values: dict[str, complex] = {k: complex(v) for k, v in params.base_parameters.items()}
    for k, v in (overrides or {}).items():
        values[getattr(k, "name", str(k))] = complex(v)
    values["t"] = complex(t)
    for name, expr in params.parsed_compound_parameters.items():
        syms = sorted(expr.free_symbols, key=lambda s: s.name)
        f = smp.lambdify(syms, expr, modules=[HELPER_FUNCTIONS, "numpy"])
        args = [values[s.name] if s.name != "t" else float(t) for s in syms]
        args = [a.real if isinstance(a, complex) and a.imag == 0 else a for a in args]
        values[name] = complex(f(*args))
    return values


def numeric_hamiltonian(prepared, t: float, overrides: dict | None = None) -> np.ndarray:
    H = prepared.obe_system.H_symbolic
    values = slot_values(prepared, t, overrides)
    syms = sorted(H.free_symbols, key=lambda s: s.name)
    f = smp.lambdify(syms, H, modules="numpy")
    full = np.asarray(f(*[values[s.name] for s in syms]), dtype=np.complex128)
    # The solver defines H by its upper triangle and Hermitian symmetry
    # (ir.py:677-879 lowers only j >= i); a symbol declared real but bound to a
    # complex runtime value makes the sympy lower triangle disagree with that.
    upper = np.triu(full)
    return upper + np.triu(full, 1).conj().T


def dense_rhs(prepared, rho: np.ndarray, t: float, overrides: dict | None = None) -> np.ndarray:
    H = numeric_hamiltonian(prepared, t, overrides)
    out = -1j * (H @ rho - rho @ H)
    for c in np.asarray(prepared.obe_system.C_array, dtype=np.complex128):
        cdc = c.conj().T @ c
        out += c @ rho @ c.conj().T - 0.5 * (cdc @ rho + rho @ cdc)
    return out


def random_density(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (m + m.conj().T).astype(np.complex128)


def chain_system(n_states: int):
    """Nearest-neighbour chain of tests/lindblad/test_rust_backend.py:489-529."""
    from centrex_tlf_amd.lindblad.systems import OBESystem

    omega, delta = smp.symbols("Omega delta", real=True)
    H = smp.zeros(n_states, n_states)
    for i in range(n_states):
        H[i, i] = delta * i
    for i in range(n_states - 1):
        H[i, i + 1] = omega / 2
        H[i + 1, i] = smp.conjugate(omega) / 2
    C = np.zeros((n_states - 1, n_states, n_states), dtype=np.complex128)
    for i in range(n_states - 1):
        C[i, i, i + 1] = np.sqrt(0.1)
    return OBESystem(H_symbolic=H, C_array=C)


CHAIN_PARAMETERS = {
    "omega0": 1.0,
    "modulation_depth": 0.5,
    "modulation_frequency": 1.9,
    "delta": 0.2,
    "Omega": "omega0*phase_modulation(t, modulation_depth, modulation_frequency)",
}
