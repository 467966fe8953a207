"""Fixed-step solvers on the GPU (fixed_dopri5, fixed_rk2, fixed_rk4;
rust/src/ode/fixed.rs) against the oracle, whose steppers
tests/test_oracle_fixed.py pins to their stability polynomials: a seeded
subset of the n=39 grid (populations, photon counts, step and RHS counts),
the public API with save points the steps must land on, and the reference's
error text.  Stage sums are written in the reference's operation order in
both, so the remaining difference is the RHS summation order (~1e-15)."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, solve_lindblad, solve_lindblad_batch, two_level_system
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39

pytestmark = pytest.mark.gpu

SOLVERS = ["fixed_dopri5", "fixed_rk2", "fixed_rk4"]
CALLS = {"fixed_dopri5": 6, "fixed_rk2": 2, "fixed_rk4": 4}


@pytest.fixture(scope="module")
def n39():
    require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    return prepared, prepared.layout.pack(rho0), t_span, weights, Oracle(prepared)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("restrict", [True, False], ids=["restricted", "full"])
def test_fixed_step_grid_subset_matches_oracle(n39, solver, restrict):
    prepared, y0, t_span, weights, oracle = n39
    rng = np.random.default_rng(20261202 + 7)
    pb = np.stack([rng.uniform(-5, 5, 6) * GAMMA, rng.uniform(0.1, 10, 6) * GAMMA], axis=1).astype(complex)
    dt = 3.3e-9  # ~620 steps over 20/Gamma, the last one shortened
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index("delta0"), prepared.slot_names.index("Omega0")]
    pops, stats, _, phot = plan.solve(y0, *t_span, solver=solver, dt=dt, output_kind="populations",
                                      weights=weights, with_integral=True, override_slots=slots, override_values=pb,
                                      restrict_support=restrict)
    kw = dict(solver=solver, dt=dt, output_when="final", parameter_slots=["delta0", "Omega0"], parameter_batch=pb)
    ref = oracle.solve_batch(y0, t_span, output="populations", **kw)
    ref_i = oracle.solve_batch(y0, t_span, output="photon_integral", integral_weights=weights, **kw)
    st = stats.cpu().numpy()
    np.testing.assert_allclose(pops.cpu().numpy(), ref["final"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(phot.cpu().numpy(), ref_i["final"][:, 0], rtol=1e-12, atol=0)
    np.testing.assert_array_equal(st[:, 0], ref["accepted"])
    np.testing.assert_array_equal(st[:, 2], ref["accepted"] * CALLS[solver])
    assert np.all(st[:, 1] == 0) and np.all(st[:, 3] == 0)


@pytest.mark.parametrize("solver", SOLVERS)
def test_fixed_step_api_lands_on_save_points(solver):
    require_gpu()
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.8, "delta": 0.3}, backend="gpu")
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    saveat = np.array([0.013, 0.25, 0.377, 0.5])
    res = solve_lindblad(prepared, rho0, (0.0, 0.5), solver=solver, dt=0.02, output="populations",
                         output_when="saveat", saveat=saveat, collect_stats=True)
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 0.5), solver=solver, dt=0.02,
                                       output="populations", output_when="saveat", saveat=saveat)
    np.testing.assert_allclose(res.t, ref["times"][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(res.values, ref["values"][0], rtol=0, atol=1e-14)
    assert res.solver_stats["accepted_steps"] == int(ref["accepted"][0])


def test_fixed_step_errors():
    require_gpu()
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.8, "delta": 0.0}, backend="gpu")
    packed = prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex))
    with pytest.raises(ValueError, match=r"^trajectory 0: fixed-step solvers require positive finite dt$"):
        solve_lindblad_batch(prepared, np.stack([packed, packed]), (0.0, 0.5), solver="fixed_rk4", dt=0.0)
    with pytest.raises(ValueError, match=r"^trajectory 0: Stopped at x=[0-9.]+\. Need more than 5 steps\.$"):
        solve_lindblad_batch(prepared, np.stack([packed, packed]), (0.0, 0.5), solver="fixed_rk2", dt=0.01,
                             maxiters=5)
    with pytest.raises(ValueError, match="Stopped at x=0.05. Need more than 5 steps"):
        solve_lindblad(prepared, np.diag([1.0, 0.0]).astype(complex), (0.0, 0.5), solver="fixed_rk2", dt=0.01,
                       maxiters=5)
