"""The oracle's terminal stop events (rust/src/ode/dopri5.rs:280-352,
tsit5.rs:274-337, fixed.rs:186-238, rust/src/lindblad/ode_impl.rs:134-153),
pinned by mathematics rather than by the formulas they restate:

* a time event g = t - t_stop lands on t_stop (60 bisections resolve the step
  to rounding), and its state is the solver's own dense output at t_stop from
  an unstopped solve with the same steps;
* undamped Rabi flopping has P_e(t) = sin^2(Omega t / 2), so population
  events have closed-form roots; threshold 0.25 separates the excited index
  (root pi / (3 Omega)) from the ground one (2 pi / (3 Omega));
* fixed-step roots are step ends, and the state there is the product of the
  steppers' stability polynomials R(hA) (see test_oracle_fixed.py);
* g = 0 at t0 fires on the first accepted step at t0 itself;
* save points past the root are dropped without an error; the root is
  pushed once.
"""

from __future__ import annotations

import numpy as np
import pytest
import sympy as smp

from centrex_tlf_amd.lindblad import normalize_stop_event, population, prepare_lindblad_problem, two_level_system
from centrex_tlf_amd.lindblad.systems import OBESystem
from oracle.lb_oracle import Oracle, OracleError

from .test_oracle_fixed import _generator, _step

OMEGA = 1.3
T = smp.Symbol("t", real=True)


def rabi_system():
    """Two levels, H = [[0, Omega/2], [Omega/2, -delta]], no decay."""
    om, de = smp.symbols("Omega delta", real=True)
    H = smp.Matrix([[0, om / 2], [smp.conjugate(om) / 2, -de]])
    return OBESystem(H_symbolic=H, C_array=np.zeros((0, 2, 2), dtype=complex), QN=[{"index": 0}, {"index": 1}],
                     description="undamped two-level system")


@pytest.fixture(scope="module")
def rabi():
    system = rabi_system()
    prepared = prepare_lindblad_problem(system, {"Omega": OMEGA, "delta": 0.0}, backend="host")
    rho0 = np.zeros((2, 2), dtype=complex)
    rho0[0, 0] = 1.0
    return prepared, Oracle(prepared), prepared.layout.pack(rho0)


@pytest.fixture(scope="module")
def damped():
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.8, "delta": 0.3}, backend="host")
    rho0 = np.zeros((2, 2), dtype=complex)
    rho0[0, 0] = 1.0
    return prepared, Oracle(prepared), prepared.layout.pack(rho0)


TOL = dict(abstol=1e-12, reltol=1e-10, dt=1e-3)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_time_event_lands_on_t_stop_with_the_dense_output_state(damped, solver):
    prepared, oracle, y0 = damped
    t_stop = 0.6180339887
    ev = normalize_stop_event(T - t_stop, prepared)
    r = oracle.solve_batch(y0, (0.0, 2.0), solver=solver, output="full", output_when="final", stop_event=ev, **TOL)
    assert abs(r["event_t"][0] - t_stop) <= 4e-16 * t_stop
    # the same steps without the event, dense output at t_stop
    ref = oracle.solve_batch(y0, (0.0, 2.0), solver=solver, output="full", output_when="saveat",
                             saveat=np.array([t_stop, 2.0]), save_start=False, **TOL)
    np.testing.assert_allclose(r["final"][0], ref["values"][0, 0], rtol=0, atol=1e-15)
    # the event step is the one that contains t_stop: steps up to it are those of the full solve
    assert 0 < r["accepted"][0] < ref["accepted"][0]
    assert r["n_times"][0] == 1
    # a wrong event (t - 2 t_stop) lands elsewhere
    ev2 = normalize_stop_event(T - 2 * t_stop, prepared)
    r2 = oracle.solve_batch(y0, (0.0, 2.0), solver=solver, output="full", output_when="final", stop_event=ev2, **TOL)
    assert abs(r2["event_t"][0] - 2 * t_stop) <= 1e-15 * 2 * t_stop


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
@pytest.mark.parametrize("index, root", [(1, np.pi / (3 * OMEGA)), (0, 2 * np.pi / (3 * OMEGA))])
def test_population_event_roots_of_rabi_flopping(rabi, solver, index, root):
    prepared, oracle, y0 = rabi
    ev = normalize_stop_event(population(index, threshold=0.25), prepared)
    r = oracle.solve_batch(y0, (0.0, 5.0), solver=solver, output="populations", output_when="final",
                           stop_event=ev, **TOL)
    assert abs(r["event_t"][0] - root) < 1e-8
    pops = r["final"][0]
    assert abs(pops[index] - 0.25) < 1e-9
    np.testing.assert_allclose(pops, [np.cos(OMEGA * root / 2) ** 2, np.sin(OMEGA * root / 2) ** 2], atol=1e-8)
    # a sum over both states is the trace (1): 1 - 0.5 never crosses
    ev_none = normalize_stop_event(population([0, 1], threshold=0.5), prepared)
    r = oracle.solve_batch(y0, (0.0, 5.0), solver=solver, output="populations", output_when="final",
                           stop_event=ev_none, **TOL)
    assert np.isnan(r["event_t"][0])


def test_expression_of_parameters(rabi):
    """g = Omega * t - pi: the root is pi / Omega, read through the parameter slot."""
    prepared, oracle, y0 = rabi
    omega = smp.Symbol("Omega", real=True)
    ev = normalize_stop_event(omega * T - smp.pi, prepared)
    r = oracle.solve_batch(y0, (0.0, 5.0), output="populations", output_when="final", stop_event=ev, **TOL)
    assert abs(r["event_t"][0] - np.pi / OMEGA) < 1e-15
    # per-trajectory overrides reach the event
    pb = np.array([[1.0], [2.0]], dtype=complex)
    r = oracle.solve_batch(y0, (0.0, 5.0), output="populations", output_when="final", stop_event=ev,
                           parameter_slots=["Omega"], parameter_batch=pb, **TOL)
    np.testing.assert_allclose(r["event_t"], [np.pi, np.pi / 2], rtol=1e-15)
    # Rabi population at a full pi pulse: all excited
    np.testing.assert_allclose(r["final"][:, 1], [1.0, 1.0], atol=1e-9)


@pytest.mark.parametrize("solver", ["fixed_rk4", "fixed_dopri5", "fixed_rk2"])
def test_fixed_step_root_is_the_step_end(damped, solver):
    prepared, oracle, y0 = damped
    h = 0.05
    ev = normalize_stop_event(T - 0.123, prepared)
    r = oracle.solve_batch(y0, (0.0, 1.0), solver=solver, dt=h, output="full", output_when="final", stop_event=ev)
    assert r["event_t"][0] == pytest.approx(0.15, abs=1e-15)
    A = _generator(oracle, y0.size)
    want = np.linalg.matrix_power(_step(A, h, solver), 3) @ y0
    np.testing.assert_allclose(r["final"][0], want, rtol=0, atol=1e-14)
    assert int(r["accepted"][0]) == 3


@pytest.mark.parametrize("solver", ["dopri5", "tsit5", "fixed_rk4"])
def test_zero_at_start_fires_on_the_first_step(damped, solver):
    prepared, oracle, y0 = damped
    ev = normalize_stop_event(T - 0.0, prepared)
    r = oracle.solve_batch(y0, (0.0, 1.0), solver=solver, dt=0.05, output="full", output_when="saveat",
                           stop_event=ev, **{k: v for k, v in TOL.items() if k != "dt"})
    assert r["event_t"][0] == 0.0 and int(r["accepted"][0]) == 1
    if solver == "fixed_rk4":
        # fixed.rs pushes the step-end state at the root's time, after the t0 point
        assert r["n_times"][0] == 2
        np.testing.assert_allclose(r["times"][0], [0.0, 0.0])
    else:
        # the root (t0, y0) is the point already pushed at t0: pushed once
        assert r["n_times"][0] == 1
        np.testing.assert_allclose(r["values"][0, 0], y0, rtol=0, atol=0)


def test_save_points_past_the_root_are_dropped(damped):
    prepared, oracle, y0 = damped
    ev = normalize_stop_event(T - 0.5, prepared)
    sv = np.array([0.1, 0.3, 0.7, 0.9])
    r = oracle.solve_batch(y0, (0.0, 1.0), output="populations", output_when="saveat", saveat=sv,
                           stop_event=ev, **TOL)
    np.testing.assert_allclose(r["times"][0], [0.0, 0.1, 0.3, 0.5], rtol=1e-15)
    ref = oracle.solve_batch(y0, (0.0, 1.0), output="populations", output_when="saveat", saveat=sv, **TOL)
    np.testing.assert_allclose(r["values"][0, :3], ref["values"][0, :3], rtol=0, atol=0)


def test_complex_event_is_an_error(damped):
    prepared, oracle, y0 = damped
    ev = normalize_stop_event(smp.I * T - 0.5, prepared)
    with pytest.raises(OracleError, match="event expression evaluated to a complex value"):
        oracle.solve_batch(y0, (0.0, 1.0), output="populations", output_when="final", stop_event=ev, **TOL)
