"""The effective-Hamiltonian path on the GPU against the oracle (pinned in
tests/test_oracle_effective.py): the RHS bitwise (same products and sums in
the same order), batched solves with per-trajectory overrides for every
output kind and both adaptive solvers, save-plan traces, a plan large enough
for several waves per trajectory, the main-path cross-check and the
field-range error."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, solve_lindblad
from centrex_tlf_amd.lindblad.effective import (
    debug_effective_lindblad_rhs,
    prepare_effective_lindblad_plan,
    solve_effective_lindblad,
    solve_effective_lindblad_batch,
    split_real_rho,
)
from oracle.lb_oracle import EffectiveOracle

from ._effective import random_plan_dict, two_level_obe, two_level_plan_dict
from ._gpu import require_gpu

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-3)


@pytest.fixture(scope="module", params=[(4, 0.15), (16, 0.02)], ids=["n4", "n16"])
def random_plan(request):
    require_gpu()
    n, density = request.param
    d = random_plan_dict(n=n, seed=7 + n, density=density)
    plan = prepare_effective_lindblad_plan(d)
    rho0 = np.zeros((n, n), dtype=complex)
    rho0[0, 0] = rho0[1, 1] = 0.5
    rho0[0, 1] = rho0[1, 0] = 0.2
    return plan, EffectiveOracle(plan), split_real_rho(rho0)


def test_rhs_is_bitwise_the_oracle(random_plan):
    plan, oracle, y0 = random_plan
    rng = np.random.default_rng(11)
    for t in (0.0, 0.9, 2.5, 3.7, 5.4):
        y = rng.normal(size=plan.real_dim)
        got = debug_effective_lindblad_rhs(plan, y, t)["dy"]
        np.testing.assert_array_equal(got, oracle.rhs(y, t))


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
@pytest.mark.parametrize("output", ["populations", "full", "selected", "photon_integral"])
def test_batch_matches_the_oracle(random_plan, solver, output):
    plan, oracle, y0 = random_plan
    n = plan.n_states
    pb = np.array([[0.4, -0.9], [1.0, -0.85], [1.7, -0.8], [0.8, -0.76]])  # x(5) = x0 + 1.75 stays on the grid
    sel = [(0, 0), (0, 1), (1, 0), (n - 1, 0)]
    weights = [((n - 1) * n + (n - 1), 0.3), (n + 1, 0.1)]
    kw = dict(solver=solver, output=output, output_indices=sel if output == "selected" else None,
              integral_weights=weights if output == "photon_integral" else None, **TOL)
    # a state output at t1 is the save plan {t1}; the integral pushes t0 and
    # every accepted step (a single pushed point integrates to zero)
    integral = output == "photon_integral"
    times, flat, width, count, stats = solve_effective_lindblad_batch(
        plan, y0, 0.0, 5.0, saveat=None if integral else [5.0], save_start=integral, parameter_slot_indices=[
            plan.slot_names.index("Omega_amp"), plan.slot_names.index("x0")],
        parameter_batch=pb, trajectory_count=4, **kw)
    ref = oracle.solve_batch(y0, (0.0, 5.0), output=output, output_when="final",
                             output_indices=kw["output_indices"], integral_weights=kw["integral_weights"],
                             parameter_slots=["Omega_amp", "x0"], parameter_batch=pb.astype(complex),
                             solver=solver, **TOL)
    got = flat.reshape(4, -1)
    want = ref["final"]
    if output == "selected":
        want = np.stack([want.real, want.imag], axis=-1).reshape(4, -1)
    # the integral is a trapezoid over accepted steps, which agree to ~1e-8
    np.testing.assert_allclose(got, want, rtol=1e-6 if integral else 0, atol=0 if integral else 1e-8)
    assert stats["accepted_steps"] == int(ref["accepted"].sum())
    if output != "photon_integral":
        np.testing.assert_allclose(times, [5.0])


def test_trace_with_save_points(random_plan):
    plan, oracle, y0 = random_plan
    sv = np.linspace(0.0, 5.0, 11)
    t, vals, width = solve_effective_lindblad(plan, y0, 0.0, 5.0, saveat=sv, output="populations", **TOL)
    ref = oracle.solve_batch(y0, (0.0, 5.0), output="populations", output_when="saveat", saveat=sv, **TOL)
    np.testing.assert_allclose(t, ref["times"][0], rtol=0, atol=0)
    np.testing.assert_allclose(vals.reshape(t.size, width), ref["values"][0], rtol=0, atol=1e-8)
    # every accepted step when no save points are given
    t, vals, width = solve_effective_lindblad(plan, y0, 0.0, 5.0, output="full", **TOL)
    ref = oracle.solve_batch(y0, (0.0, 5.0), output="full", output_when="saveat", **TOL)
    assert t.size == ref["n_times"][0]
    np.testing.assert_allclose(vals.reshape(t.size, width)[-1], ref["values"][0][-1], rtol=0, atol=1e-8)


def test_two_level_plan_matches_the_main_gpu_path():
    require_gpu()
    plan = prepare_effective_lindblad_plan(two_level_plan_dict())
    rho0 = np.diag([1.0, 0.0]).astype(complex)
    t, vals, width = solve_effective_lindblad(plan, split_real_rho(rho0), 0.0, 5.0, saveat=[5.0],
                                              save_start=False, output="populations", abstol=1e-11, reltol=1e-9,
                                              dt=1e-3)
    prepared = prepare_lindblad_problem(two_level_obe(), {"Omega": 1.1, "delta": 0.2, "x0": -1.0, "v": 0.5})
    res = solve_lindblad(prepared, rho0, (0.0, 5.0), output="populations", output_when="final", abstol=1e-11,
                         reltol=1e-9, dt=1e-3)
    np.testing.assert_allclose(vals, res.values, rtol=0, atol=1e-8)


def test_field_outside_the_grid():
    require_gpu()
    plan = prepare_effective_lindblad_plan(two_level_plan_dict())
    y0 = split_real_rho(np.diag([1.0, 0.0]).astype(complex))
    with pytest.raises(ValueError, match=r"^trajectory 0: field coordinate 2\.0\d{5}e0 is above the operator grid "
                                         r"maximum 2\.000000e0$"):
        solve_effective_lindblad_batch(plan, y0, 0.0, 7.0, 1e-10, 1e-8, 1e-3, saveat=[7.0], save_start=False)
