"""Pin the oracle's starting-step choice (hinit, rust/src/ode/common.rs:72-118)
against Hairer's published HINIT (Hairer, Norsett & Wanner, Solving ODEs I,
sec. II.4, and the HINIT routine of dopri5.f), written here in numpy from that
description over the textbook dense Lindblad right-hand side of tests/_dense.py
(not the oracle's RHS).

The cases reach every branch: the regular estimate, the d0/d1 < 1e-10 fallback
(h0 = 1e-6), the der12 <= 1e-15 fallback (a state that does not move), and the
clip to the span; the two-level system and the n = 39 TlF system.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, two_level_system
from oracle.lb_oracle import Oracle

from ._dense import dense_rhs, random_density
from .systems import GAMMA, tlf_n39


def hairer_hinit(f, x, y, x_end, atol, rtol, order=5):
    """HINIT: h from ||y||, ||f0|| and a finite-difference second derivative."""
    posneg = 1.0 if x_end - x >= 0 else -1.0
    hmax = abs(x_end - x)
    sk = atol + rtol * np.abs(y)
    f0 = f(x, y)
    dnf = float(np.sum((f0 / sk) ** 2))
    dny = float(np.sum((y / sk) ** 2))
    h = 1.0e-6 if (dnf <= 1.0e-10 or dny <= 1.0e-10) else np.sqrt(dny / dnf) * 0.01
    h = posneg * min(h, hmax)
    f1 = f(x + h, y + h * f0)
    der2 = np.sqrt(np.sum(((f1 - f0) / sk) ** 2)) / h
    der12 = max(abs(der2), np.sqrt(dnf))
    h1 = max(1.0e-6, abs(h) * 1.0e-3) if der12 <= 1.0e-15 else (0.01 / der12) ** (1.0 / order)
    return posneg * min(100 * abs(h), h1, hmax)


def packed_rhs(prepared, overrides=None):
    layout = prepared.layout

    def f(t, y):
        return layout.pack(dense_rhs(prepared, layout.unpack(y), t, overrides))

    return f


def two_level(omega, delta, decay):
    system = two_level_system(decay=decay, omega_name="Omega", delta_name="delta")
    return prepare_lindblad_problem(system, {"Omega": omega, "delta": delta}, backend="host")


@pytest.mark.parametrize("omega,delta,decay,t_span", [
    (0.6, 0.0, 0.3, (0.0, 0.5)),
    (2.0, -1.1, 0.05, (0.3, 10.0)),
    (0.0, 0.0, 0.3, (0.0, 4.0)),        # decay only
    (5.0, 0.4, 1.0, (0.0, 1e-7)),       # clipped to the span
])
def test_two_level_hinit_is_hairers(omega, delta, decay, t_span):
    prepared = two_level(omega, delta, decay)
    oracle = Oracle(prepared)
    for seed in range(4):
        rho = np.diag([1.0, 0.0]).astype(complex) if seed == 0 else random_density(2, seed)
        y = prepared.layout.pack(rho)
        for atol, rtol in [(1e-10, 1e-8), (1e-6, 1e-3)]:
            want = hairer_hinit(packed_rhs(prepared), t_span[0], y, t_span[1], atol, rtol)
            got = oracle.hinit(y, t_span, atol, rtol)
            assert got == pytest.approx(want, rel=1e-12), (seed, atol)


def test_fallback_branches():
    # no drive and the ground state: f = 0, so d1 < 1e-10 and der12 = 0 -> h = 1e-6
    prepared = two_level(0.0, 0.0, 0.3)
    y = prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex))
    got = Oracle(prepared).hinit(y, (0.0, 1.0), 1e-10, 1e-8)
    assert got == hairer_hinit(packed_rhs(prepared), 0.0, y, 1.0, 1e-10, 1e-8) == 1e-6
    # a zero state: d0 < 1e-10 -> h0 = 1e-6, then der2 = 0 and |f0| = 0 -> 1e-6
    prepared = two_level(1.0, 0.0, 0.3)
    y = np.zeros(4)
    assert Oracle(prepared).hinit(y, (0.0, 1.0), 1e-10, 1e-8) == 1e-6


def test_n39_hinit_is_hairers():
    system, params, rho0, t_span, _ = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    oracle = Oracle(prepared)
    y = prepared.layout.pack(rho0)
    for delta, om in [(0.0, 1.0), (2.5, 0.3), (-4.0, 7.0)]:
        ov = {"delta0": delta * GAMMA, "Omega0": om * GAMMA}
        for t0 in (0.0, 8.0 / GAMMA):
            want = hairer_hinit(packed_rhs(prepared, ov), t0, y, t_span[1], 1e-10, 1e-8)
            got = oracle.hinit(y, (t0, t_span[1]), 1e-10, 1e-8, ov)
            assert got == pytest.approx(want, rel=1e-10), (delta, om, t0)
