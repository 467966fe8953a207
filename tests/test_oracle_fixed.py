"""The oracle's fixed-step solvers (rust/src/ode/fixed.rs:133-249), pinned by
mathematics.  For a time-independent plan the right-hand side is linear,
dy/dt = A y, so one explicit Runge-Kutta step of size h maps y to R(hA) y,
with R the stepper's stability polynomial

    R(Z) = I + sum_k (b^T T^k 1) Z^(k+1)

built here from the published Butcher tableau (T, b) -- independent of the
stage loop the oracle runs, so a wrong coefficient, sign or stage fails.  A is
assembled column by column from the oracle's RHS, which test_lowering.py pins
against the dense Lindblad equation.  Step placement (dt, the shortened last
step, landing on save-plan points) and the reference's error paths are
checked the same way."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, two_level_system
from oracle.lb_oracle import Oracle, OracleError

from ._dense import chain_system

TABLEAUS = {
    "fixed_rk2": (np.array([[0.0, 0.0], [0.5, 0.0]]), np.array([0.0, 1.0])),
    "fixed_rk4": (np.array([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]], dtype=float),
                  np.array([1 / 6, 1 / 3, 1 / 3, 1 / 6])),
    "fixed_dopri5": (np.array([
        [0, 0, 0, 0, 0, 0],
        [1 / 5, 0, 0, 0, 0, 0],
        [3 / 40, 9 / 40, 0, 0, 0, 0],
        [44 / 45, -56 / 15, 32 / 9, 0, 0, 0],
        [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729, 0, 0],
        [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656, 0]]),
        np.array([35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84])),
}
RHS_CALLS = {"fixed_rk2": 2, "fixed_rk4": 4, "fixed_dopri5": 6}


def _systems():
    two = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    p2 = prepare_lindblad_problem(two, {"Omega": 1.3, "delta": 0.4}, backend="host")
    y2 = np.zeros((2, 2), dtype=complex)
    y2[0, 0] = 1.0
    p3 = prepare_lindblad_problem(chain_system(3), {"Omega": 0.8, "delta": 0.2}, backend="host")
    y3 = np.zeros((3, 3), dtype=complex)
    y3[0, 0] = 0.7
    y3[1, 1] = 0.3
    y3[0, 1] = y3[1, 0] = 0.2
    return [(p2, p2.layout.pack(y2)), (p3, p3.layout.pack(y3))]


def _generator(oracle, dim):
    return np.stack([oracle.rhs(np.eye(dim)[i], 0.0) for i in range(dim)], axis=1)


def _step(A, h, solver):
    T, b = TABLEAUS[solver]
    Z = h * A
    R = np.eye(A.shape[0])
    Tk1 = np.ones(len(b))
    Zk = Z.copy()
    for _ in range(len(b)):
        R = R + (b @ Tk1) * Zk
        Tk1 = T @ Tk1
        Zk = Zk @ Z
    return R


@pytest.fixture(scope="module", params=[0, 1], ids=["two_level", "chain3"])
def system(request):
    prepared, y0 = _systems()[request.param]
    oracle = Oracle(prepared)
    return prepared, oracle, y0, _generator(oracle, y0.size)


@pytest.mark.parametrize("solver", list(TABLEAUS))
def test_steps_are_the_stability_polynomial(system, solver):
    prepared, oracle, y0, A = system
    h = 0.05
    # three full steps and a half step: dt = h over [0, 3.5 h]
    r = oracle.solve_batch(y0, (0.0, 3.5 * h), solver=solver, dt=h, output="full", output_when="final")
    want = _step(A, 0.5 * h, solver) @ np.linalg.matrix_power(_step(A, h, solver), 3) @ y0
    np.testing.assert_allclose(r["final"][0], want, rtol=0, atol=1e-14)
    assert int(r["accepted"][0]) == 4 and int(r["rhs_calls"][0]) == 4 * RHS_CALLS[solver]
    # a mistake in the stepper shows: another stepper's polynomial differs at O(h^3)
    other = "fixed_rk2" if solver != "fixed_rk2" else "fixed_rk4"
    assert np.abs(r["final"][0] - _step(A, 0.5 * h, other) @ np.linalg.matrix_power(_step(A, h, other), 3) @ y0).max() > 1e-9


@pytest.mark.parametrize("solver", ["fixed_rk4", "fixed_dopri5"])
def test_steps_land_on_save_points(system, solver):
    prepared, oracle, y0, A = system
    h = 0.05
    r = oracle.solve_batch(y0, (0.0, 2.0 * h), solver=solver, dt=h, output="full", output_when="saveat",
                           saveat=np.array([0.3 * h, 1.7 * h, 2.0 * h]))
    np.testing.assert_allclose(r["times"][0], [0.0, 0.3 * h, 1.7 * h, 2.0 * h], rtol=1e-15)
    # steps 0.3h, h, 0.4h, 0.3h
    s03, s1, s04 = _step(A, 0.3 * h, solver), _step(A, h, solver), _step(A, 0.4 * h, solver)
    pts = [y0, s03 @ y0, s04 @ s1 @ s03 @ y0, _step(A, 0.3 * h, solver) @ s04 @ s1 @ s03 @ y0]
    np.testing.assert_allclose(r["values"][0], np.stack(pts), rtol=0, atol=1e-14)
    assert int(r["accepted"][0]) == 4


def test_fixed_step_errors(system):
    prepared, oracle, y0, A = system
    with pytest.raises(OracleError, match="fixed-step solvers require positive finite dt"):
        oracle.solve_batch(y0, (0.0, 1.0), solver="fixed_rk4", dt=0.0, output="full", output_when="final")
    with pytest.raises(OracleError, match="Need more than 3 steps"):
        oracle.solve_batch(y0, (0.0, 1.0), solver="fixed_rk4", dt=0.1, maxiters=3, output="full",
                           output_when="final")
    with pytest.raises(OracleError, match="only forward integration"):
        oracle.solve_batch(y0, (1.0, 0.0), solver="fixed_rk2", dt=0.1, output="full", output_when="final")
