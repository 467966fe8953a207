"""Terminal stop events on the GPU against the oracle (whose event logic
tests/test_oracle_events.py pins by closed forms): population thresholds,
expressions of t and of base and time-dependent parameters, adaptive and
fixed-step solvers, save plans, restricted plans, and the API's result
layout (``t`` holds each trajectory's root, ``metadata`` the flags).

Step sizes agree with the oracle to ~1e-8 relative (not bitwise, see the
adaptive parity tests), so roots of population events agree to the solution
accuracy (1e-7 relative); roots of expression events in t alone are resolved
by the 60 bisections to rounding."""

from __future__ import annotations

import math

import numpy as np
import pytest
import sympy as smp

from centrex_tlf_amd.lindblad import (
    grid_scan,
    normalize_stop_event,
    population,
    prepare_lindblad_problem,
    solve_lindblad,
    solve_lindblad_batch,
    two_level_system,
)
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39

pytestmark = pytest.mark.gpu

T = smp.Symbol("t", real=True)
TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-9)


@pytest.fixture(scope="module")
def n39():
    require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    rng = np.random.default_rng(20261202 + 11)
    pb = np.stack([rng.uniform(-5, 5, 12) * GAMMA, rng.uniform(0.1, 10, 12) * GAMMA], axis=1).astype(complex)
    return prepared, prepared.layout.pack(rho0), t_span, weights, Oracle(prepared), pb


def _raw_events():
    excited = [36, 37, 38]
    omega0 = smp.Symbol("Omega0", real=True)
    env = smp.Symbol("Ω0", real=True)
    return {
        # total excited population reaching 2 %: fires for the stronger drives only
        "excited_population": population(excited, threshold=0.02),
        # a base parameter: root 10 / Omega0 inside the span for Omega0 > Gamma / 2
        "omega_time": omega0 * T - 10.0,
        # the time-dependent Gaussian envelope at half its peak: t_center - sigma sqrt(2 ln 2)
        # (scaled: the envelope carries a ~1e-16 relative imaginary part from the
        # complex power of a negative base, which in units of Omega0 ~ 1e7 would
        # exceed the 1e-12 limit of a real event value)
        "half_envelope": env / omega0 - 0.5,
    }


def _events(prepared):
    return {k: normalize_stop_event(v, prepared) for k, v in _raw_events().items()}


@pytest.mark.parametrize("name", ["excited_population", "omega_time", "half_envelope"])
@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_adaptive_events_match_the_oracle(n39, name, solver):
    prepared, y0, t_span, weights, oracle, pb = n39
    ev = _events(prepared)[name]
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index("delta0"), prepared.slot_names.index("Omega0")]
    pops, stats, _, phot, event_t = plan.solve(
        y0, *t_span, solver=solver, output_kind="populations", weights=weights, with_integral=True,
        override_slots=slots, override_values=pb, stop_event=ev, **TOL)
    kw = dict(solver=solver, output_when="final", parameter_slots=["delta0", "Omega0"], parameter_batch=pb,
              stop_event=ev, **TOL)
    ref = oracle.solve_batch(y0, t_span, output="populations", **kw)
    ref_i = oracle.solve_batch(y0, t_span, output="photon_integral", integral_weights=weights, **kw)
    got_t = event_t.cpu().numpy()
    st = stats.cpu().numpy()
    assert np.all(st[:, 3] == 0)
    np.testing.assert_array_equal(np.isfinite(got_t), np.isfinite(ref["event_t"]))
    fired = np.isfinite(got_t)
    assert 0 < fired.sum() <= fired.size
    np.testing.assert_allclose(got_t[fired], ref["event_t"][fired], rtol=1e-7 if name == "excited_population" else 1e-13)
    np.testing.assert_allclose(pops.cpu().numpy(), ref["final"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(phot.cpu().numpy(), ref_i["final"][:, 0], rtol=1e-6, atol=1e-12)
    np.testing.assert_array_equal(st[:, 0], ref["accepted"])
    if name == "omega_time":
        np.testing.assert_allclose(got_t[fired], 10.0 / pb[fired, 1].real, rtol=1e-14)
    if name == "half_envelope":
        root = 10.0 / GAMMA - 3.0 / GAMMA * math.sqrt(2.0 * math.log(2.0))
        np.testing.assert_allclose(got_t, root, rtol=1e-12)


@pytest.mark.parametrize("solver", ["fixed_rk4", "fixed_dopri5"])
def test_fixed_step_events_match_the_oracle(n39, solver):
    prepared, y0, t_span, weights, oracle, pb = n39
    ev = _events(prepared)["excited_population"]
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index("delta0"), prepared.slot_names.index("Omega0")]
    dt = 5e-9
    pops, stats, _, event_t = plan.solve(y0, *t_span, solver=solver, dt=dt, output_kind="populations",
                                         override_slots=slots, override_values=pb, stop_event=ev)
    ref = oracle.solve_batch(y0, t_span, solver=solver, dt=dt, output="populations", output_when="final",
                             parameter_slots=["delta0", "Omega0"], parameter_batch=pb, stop_event=ev)
    got_t = event_t.cpu().numpy()
    np.testing.assert_array_equal(np.isfinite(got_t), np.isfinite(ref["event_t"]))
    fired = np.isfinite(got_t)
    assert fired.any()
    # the roots are step ends: the same step of the same grid
    np.testing.assert_allclose(got_t[fired], ref["event_t"][fired], rtol=1e-15)
    np.testing.assert_allclose(pops.cpu().numpy(), ref["final"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(stats.cpu().numpy()[:, 0], ref["accepted"])


def test_restricted_plan_population_outside_the_support(n39):
    """J=1 levels are spectators (exactly zero in the restricted plan): a
    threshold 0 event on one of them is zero at t0 and fires at t0."""
    prepared, y0, t_span, weights, oracle, pb = n39
    ev = normalize_stop_event(population(4, threshold=0.0), prepared)
    plan = prepared.device_plan()
    pops, stats, _, event_t = plan.solve(y0, *t_span, output_kind="populations", n_traj=3, stop_event=ev, **TOL)
    ref = oracle.solve_batch(y0, t_span, output="populations", output_when="final", stop_event=ev, **TOL)
    np.testing.assert_array_equal(event_t.cpu().numpy(), [0.0, 0.0, 0.0])
    np.testing.assert_allclose(pops.cpu().numpy()[0], ref["final"][0], rtol=0, atol=0)
    assert stats.cpu().numpy()[0, 0] == ref["accepted"][0] == 1


def test_batch_api_result_layout(n39):
    prepared, y0, t_span, weights, oracle, pb = n39
    ev = _raw_events()["omega_time"]
    axes = {"Omega0": np.array([0.25, 1.0, 4.0]) * GAMMA}
    res = grid_scan(prepared, y0, t_span, scan=axes, output="populations", stop_event=ev,
                    collect_stats=True, **TOL)
    fired = res.metadata["event_triggered"]
    np.testing.assert_array_equal(fired, [False, True, True])
    want = np.where(fired, 10.0 / axes["Omega0"], t_span[1])
    np.testing.assert_allclose(res.t, want, rtol=1e-14)
    np.testing.assert_allclose(res.metadata["event_times"], want, rtol=1e-14)
    st = res.solver_stats
    assert st["event_triggered"] is True and st["event_count"] == 2
    assert st["event_triggered_by_trajectory"] == [False, True, True]
    batch = np.stack([y0, y0])
    res = solve_lindblad_batch(prepared, batch, t_span, output="photon_integral", integral_weights=weights,
                               stop_event=_raw_events()["half_envelope"], **TOL)
    assert res.values.shape == (2, 1) and np.all(res.metadata["event_triggered"])
    with pytest.raises(ValueError, match="stop_event is only supported with output_when='final'"):
        solve_lindblad_batch(prepared, batch, t_span, output_when="saveat", saveat=[1e-6], stop_event=ev)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5", "fixed_rk4"])
def test_single_solve_trace_ends_at_the_root(solver):
    require_gpu()
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.8, "delta": 0.3}, backend="gpu")
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    ev = population(1, threshold=0.05)
    sv = np.array([0.1, 0.3, 0.5, 0.7, 0.9, 1.2])
    kw = dict(abstol=1e-10, reltol=1e-8, dt=0.01)
    res = solve_lindblad(prepared, rho0, (0.0, 1.2), solver=solver, output="populations",
                         output_when="saveat", saveat=sv, stop_event=ev, collect_stats=True, **kw)
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 1.2), solver=solver, output="populations",
                                       output_when="saveat", saveat=sv, stop_event=normalize_stop_event(ev, prepared),
                                       **kw)
    np.testing.assert_allclose(res.t, ref["times"][0], rtol=1e-9)
    np.testing.assert_allclose(res.values, ref["values"][0], rtol=0, atol=1e-9)
    assert res.t[-1] < 1.2 and res.t[-1] == pytest.approx(ref["event_t"][0], rel=1e-9)
    assert res.solver_stats["event_triggered"] and res.solver_stats["event_index"] == 0
    assert res.solver_stats["event_name"] == "population"
    # every accepted step pushed, then the root
    res = solve_lindblad(prepared, rho0, (0.0, 1.2), solver=solver, output="full", stop_event=ev, **kw)
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 1.2), solver=solver, output="full",
                                       output_when="saveat", stop_event=normalize_stop_event(ev, prepared), **kw)
    # accepted-step times agree as the step sizes do (summation order in the
    # error estimate: ~1e-8 relative for Dopri5, ~1e-6 for Tsit5, whose
    # estimate cancels larger coefficients); the root itself is resolved
    assert res.t.size == ref["n_times"][0]
    step_tol = 1e-5 if solver == "tsit5" else 1e-7
    np.testing.assert_allclose(res.t, ref["times"][0], rtol=step_tol)
    np.testing.assert_allclose(res.packed_y, ref["values"][0], rtol=0, atol=step_tol)
    assert res.t[-1] == pytest.approx(ref["times"][0][-1], rel=1e-9)
    np.testing.assert_allclose(res.packed_y[-1], ref["values"][0][-1], rtol=0, atol=1e-9)


def test_complex_event_error():
    require_gpu()
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.8, "delta": 0.3}, backend="gpu")
    packed = prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex))
    with pytest.raises(ValueError, match=r"^trajectory 0: event expression evaluated to a complex value$"):
        solve_lindblad_batch(prepared, np.stack([packed, packed]), (0.0, 1.0), stop_event=smp.I * T - 0.5)
