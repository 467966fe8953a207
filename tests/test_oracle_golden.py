"""Pin the CPU oracle against the reference's own known answers.

Every number compared here is copied from a file the reference ships
(tests/golden/two_level_known_answers.json names the source of each).  A
plausible slip in the oracle -- a wrong tableau entry, a dropped decay term,
a controller off by one, a trapezoid on the wrong points -- changes the step
counts or the integrals and fails one of these.
"""

from __future__ import annotations

import csv
import json

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, two_level_system
from oracle.lb_oracle import Oracle, OracleError


def _two_level(omega: float, decay: float = 0.3):
    system = two_level_system(decay=decay, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": omega, "delta": 0.0}, backend="host")
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    return prepared, Oracle(prepared), prepared.layout.pack(rho0)


COMMON = dict(abstol=1e-10, reltol=1e-8, dt=1e-3, output="photon_integral", output_when="final",
              integral_weights=[(1, 0.3)])


@pytest.fixture(scope="module")
def known(golden_dir):
    return json.loads((golden_dir / "two_level_known_answers.json").read_text())


def test_single_photon_integral_bitwise(known):
    ka = known["single_photon_integral"]
    _, oracle, y0 = _two_level(ka["omega"])
    r = oracle.solve_batch(y0, tuple(ka["t_span"]), **COMMON)
    assert r["final"][0, 0] == ka["value"]
    assert int(r["accepted"][0]) == ka["accepted_steps"]
    assert int(r["rejected"][0]) == ka["rejected_steps"]
    assert int(r["rhs_calls"][0]) == ka["rhs_calls"]


def test_grid_step_totals(known):
    ka = known["grid_25"]
    _, oracle, y0 = _two_level(ka["omega"])
    d = np.linspace(*ka["delta_linspace"][:2], ka["delta_linspace"][2])
    r = oracle.solve_batch(y0, tuple(ka["t_span"]), parameter_slots=["delta"],
                           parameter_batch=d[:, None].astype(complex), threads=4, **COMMON)
    assert int(r["accepted"].sum()) == ka["accepted_steps"]
    assert int(r["rhs_calls"].sum()) == ka["rhs_calls"]


def test_fixed_step_report_counts(known):
    single, grid = known["fixed_step_single"], known["fixed_step_grid_201"]
    _, oracle, y0 = _two_level(single["omega"])
    r = oracle.solve_batch(y0, tuple(single["t_span"]), **COMMON)
    assert (int(r["accepted"][0]), int(r["rhs_calls"][0])) == (single["accepted_steps"], single["rhs_calls"])
    d = np.linspace(*grid["delta_linspace"][:2], grid["delta_linspace"][2])
    r = oracle.solve_batch(y0, tuple(grid["t_span"]), parameter_slots=["delta"],
                           parameter_batch=d[:, None].astype(complex), **COMMON)
    assert int(r["accepted"].sum()) == grid["accepted_steps"]
    assert int(r["rhs_calls"].sum()) == grid["rhs_calls"]


def test_integral_scan_three_ways(known, golden_dir):
    ka = known["integral_scan_101"]
    _, oracle, y0 = _two_level(ka["omega"])
    rows = list(csv.DictReader((golden_dir / ka["csv"]).open()))
    d = np.asarray([float(r["detuning"]) for r in rows])
    assert np.allclose(d, np.linspace(-2.0, 2.0, 101))
    pb = d[:, None].astype(complex)
    base = dict(abstol=1e-10, reltol=1e-8, dt=1e-3, parameter_slots=["delta"], parameter_batch=pb)
    # trapezoid over accepted steps (the reference's photon count)
    r = oracle.solve_batch(y0, (0.0, 4.0), output="photon_integral", output_when="final",
                           integral_weights=[(1, 0.3)], **base)
    want = np.asarray([float(r_["in_solver_final"]) for r_ in rows])
    np.testing.assert_allclose(r["final"][:, 0], want, rtol=0, atol=1e-11)
    # Hermite dense output at 401 points, trapezoid afterwards
    saveat = np.linspace(0.0, 4.0, ka["save_points"])
    r = oracle.solve_batch(y0, (0.0, 4.0), output="populations", output_when="saveat",
                           saveat=saveat, **base)
    post = np.trapezoid(0.3 * r["values"][:, :, 1], x=saveat, axis=1)
    np.testing.assert_allclose(post, [float(r_["post_solve_trapezoid"]) for r_ in rows], rtol=0, atol=1e-14)
    # running integral traced at the save points
    r = oracle.solve_batch(y0, (0.0, 4.0), output="photon_integral", output_when="saveat",
                           saveat=saveat, integral_weights=[(1, 0.3)], **base)
    np.testing.assert_allclose(r["values"][:, -1, 0], [float(r_["in_solver_trace_final"]) for r_ in rows],
                               rtol=0, atol=1e-14)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_undamped_rabi_closed_form(solver):
    """rho_11(t) = (W^2/(W^2+d^2)) sin^2(sqrt(W^2+d^2) t / 2) without decay."""
    system = two_level_system(decay=0.0, omega_name="Omega", delta_name="delta")
    omega, delta = 1.3, 0.7
    from centrex_tlf_amd.lindblad.systems import OBESystem

    system = OBESystem(system.H_symbolic, np.zeros((0, 2, 2), dtype=complex))
    prepared = prepare_lindblad_problem(system, {"Omega": omega, "delta": delta}, backend="host")
    oracle = Oracle(prepared)
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    saveat = np.linspace(0.0, 6.0, 13)
    r = oracle.solve_batch(prepared.layout.pack(rho0), (0.0, 6.0), solver=solver, abstol=1e-12,
                           reltol=1e-11, dt=1e-3, output="populations", output_when="saveat",
                           saveat=saveat)
    w = np.sqrt(omega**2 + delta**2)
    expect = omega**2 / w**2 * np.sin(w * saveat / 2.0) ** 2
    np.testing.assert_allclose(r["values"][0, :, 1], expect, atol=1e-9)


def test_damped_steady_state():
    """Long-time excited population of a driven decaying two-level atom: s/(2(1+s)), s = 2W^2/(G^2+4d^2)."""
    omega, delta, gamma = 0.9, 0.4, 0.3
    system = two_level_system(decay=gamma, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": omega, "delta": delta}, backend="host")
    oracle = Oracle(prepared)
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    r = oracle.solve_batch(prepared.layout.pack(rho0), (0.0, 400.0), abstol=1e-12, reltol=1e-10,
                           dt=1e-3, output="populations", output_when="final")
    s = 2 * omega**2 / (gamma**2 + 4 * delta**2)
    np.testing.assert_allclose(r["final"][0, 1], s / (2 * (1 + s)), atol=1e-8)


def test_tsit5_matches_dopri5():
    _, oracle, y0 = _two_level(0.75)
    d = np.linspace(-2.0, 2.0, 9)[:, None].astype(complex)
    kw = dict(abstol=1e-10, reltol=1e-8, dt=1e-3, parameter_slots=["delta"], parameter_batch=d,
              output="populations", output_when="final")
    a = oracle.solve_batch(y0, (0.0, 4.0), solver="dopri5", **kw)["final"]
    b = oracle.solve_batch(y0, (0.0, 4.0), solver="tsit5", **kw)["final"]
    np.testing.assert_allclose(a, b, atol=5e-10, rtol=5e-8)


def test_error_paths():
    _, oracle, y0 = _two_level(0.6)
    with pytest.raises(OracleError, match="trajectory 0: Stopped at x=.*Need more than"):
        oracle.solve_batch(y0, (0.0, 0.5), maxiters=1, **COMMON)
    with pytest.raises(OracleError, match="only forward integration"):
        oracle.solve_batch(y0, (0.5, 0.0), **COMMON)
