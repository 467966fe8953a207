"""The oracle's effective-Hamiltonian path (rust/src/effective_lindblad/rhs.rs:47-138,
plan.rs:18-40, solver.rs:78-281), pinned by mathematics:

* with an internal Hamiltonian linear in the field, linear interpolation is
  exact, so the RHS at any t equals the dense split-real Lindblad
  superoperator at x(t) applied to y -- built here from kron products,
  independently of the operator tables, the interval lookup and the scales;
* at grid points a pchip plan reproduces the tabulated operator exactly,
  and inside interior intervals it equals scipy's PchipInterpolator (the
  same harmonic-mean slopes);
* solves agree with the main Lindblad path on the same physics;
* a field leaving the grid fails with the reference's message.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.effective import EffectivePlan, split_real_rho
from oracle.lb_oracle import EffectiveOracle, Oracle, OracleError

from ._effective import dense_rhs, random_plan_dict, two_level_obe, two_level_operators, two_level_plan_dict


@pytest.fixture(scope="module")
def linear_plan():
    plan = EffectivePlan(two_level_plan_dict(), device=False)
    return plan, EffectiveOracle(plan)


def test_linear_plan_rhs_is_the_dense_lindblad_equation(linear_plan):
    plan, oracle = linear_plan
    rng = np.random.default_rng(3)
    for t in [0.0, 0.37, 1.0, 2.9, 5.2]:  # x = -1 + t/2 crosses grid points and interiors
        y = rng.normal(size=8)
        np.testing.assert_allclose(oracle.rhs(y, t), dense_rhs(y, t), rtol=0, atol=1e-14)


def test_pchip_plan_matches_knots_and_scipy_inside():
    grid = np.array([-2.0, -1.3, -0.4, 0.2, 1.1, 2.0])
    plan = EffectivePlan(two_level_plan_dict(grid=grid, interpolation="pchip", quadratic=True, x0=0.0, v=1.0),
                         device=False)
    oracle = EffectiveOracle(plan)
    comb, _, _ = two_level_operators(grid, quadratic=True)
    y = np.random.default_rng(5).normal(size=8)
    # at the knots (x = t): the tabulated operator
    for k in (1, 2, 3):
        np.testing.assert_allclose(oracle.rhs(y, grid[k]), dense_rhs(y, grid[k], x0=0.0, v=1.0, quadratic=True),
                                   rtol=0, atol=1e-13)
    # inside interior intervals: scipy's pchip of each operator entry
    stack = np.stack(comb)  # (G, 8, 8)
    pch = PchipInterpolator(grid, stack, axis=0)
    for x in (-0.8, -0.1, 0.6):
        ref = dense_rhs(y, x, x0=0.0, v=1.0, quadratic=True) - stack[0] @ y * 0.0
        want = pch(x) @ y + (dense_rhs(y, x, x0=0.0, v=1.0) - dense_rhs(y, x, x0=0.0, v=1.0, omega=0.0, delta=0.0))
        np.testing.assert_allclose(oracle.rhs(y, x), want, rtol=1e-12, atol=1e-13)
        assert np.abs(oracle.rhs(y, x) - ref).max() > 1e-6  # pchip is not the exact quadratic


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_solve_matches_the_main_lindblad_path(linear_plan, solver):
    plan, oracle = linear_plan
    rho0 = np.diag([1.0, 0.0]).astype(complex)
    kw = dict(solver=solver, abstol=1e-11, reltol=1e-9, dt=1e-3)
    r = oracle.solve_batch(split_real_rho(rho0), (0.0, 5.0), output="populations", output_when="final", **kw)
    prepared = prepare_lindblad_problem(two_level_obe(), {"Omega": 1.1, "delta": 0.2, "x0": -1.0, "v": 0.5},
                                        backend="host")
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 5.0), output="populations",
                                       output_when="final", **kw)
    np.testing.assert_allclose(r["final"][0], ref["final"][0], rtol=0, atol=1e-8)
    # the integral of the excited population (real-vector index 1*2+1 = 3)
    ri = oracle.solve_batch(split_real_rho(rho0), (0.0, 5.0), output="photon_integral", output_when="final",
                            integral_weights=[(3, 0.3)], **kw)
    refi = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 5.0), output="photon_integral",
                                        output_when="final", integral_weights=[(1, 0.3)], **kw)
    # the trapezoid runs over each path's own accepted steps (the split-real
    # error norm counts both off-diagonal entries), so it agrees to the
    # quadrature error, not the solution accuracy
    np.testing.assert_allclose(ri["final"][0], refi["final"][0], rtol=1e-4)


def test_field_outside_the_grid_is_an_error(linear_plan):
    plan, oracle = linear_plan
    y0 = split_real_rho(np.diag([1.0, 0.0]).astype(complex))
    # x(t) = -1 + t/2 leaves [-2, 2] at t = 6
    with pytest.raises(OracleError, match=r"trajectory 0: field coordinate 2\.0\d{5}e0 is above the operator grid "
                                          r"maximum 2\.000000e0"):
        oracle.solve_batch(y0, (0.0, 7.0), output="populations", output_when="final", abstol=1e-10, reltol=1e-8,
                           dt=1e-3)


def test_random_plan_tables_and_overrides():
    plan = EffectivePlan(random_plan_dict(n=3, seed=2), device=False)
    oracle = EffectiveOracle(plan)
    assert plan.real_dim == 18 and all(op["nnz"] > 0 for op in plan.operators)
    y0 = split_real_rho(np.diag([0.5, 0.5, 0.0]).astype(complex))
    pb = np.array([[0.5], [1.5]], dtype=complex)
    r = oracle.solve_batch(y0, (0.0, 5.0), output="full", output_when="final", abstol=1e-10, reltol=1e-8,
                           dt=1e-3, parameter_slots=["Omega_amp"], parameter_batch=pb)
    assert np.all(np.isfinite(r["final"])) and np.abs(r["final"][0] - r["final"][1]).max() > 1e-6
