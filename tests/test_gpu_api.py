"""The public API on the GPU against the reference's golden vectors and the oracle.

Covers solve_lindblad (final, integral, every-step trace, saveat with dense
output), solve_lindblad_batch / initial_condition_scan / parameter_scan /
grid_scan result layouts, selected and rate outputs, and the reference's
per-trajectory error messages.  Tolerances: golden integrals 1e-11 to 1e-13,
TlF results 1e-7 (the populations bar), saveat points 1e-10, per-step traces
1e-8 (their times follow the step-size sequence, which agrees to ~1e-8
relative because the error estimate cancels O(1) stage sums).
"""

from __future__ import annotations

import csv
import json
import re

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import (
    grid_scan,
    initial_condition_scan,
    parameter_scan,
    prepare_lindblad_problem,
    solve_lindblad,
    solve_lindblad_batch,
    two_level_system,
)
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-3)


def _two_level(omega=0.6):
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": omega, "delta": 0.0}, backend="gpu")
    rho0 = np.zeros((2, 2), dtype=np.complex128)
    rho0[0, 0] = 1.0
    return prepared, rho0


def test_solve_lindblad_golden_photon_integral(golden_dir):
    require_gpu()
    known = json.loads((golden_dir / "two_level_known_answers.json").read_text())["single_photon_integral"]
    prepared, rho0 = _two_level(0.6)
    res = solve_lindblad(prepared, rho0, (0.0, 0.5), output="photon_integral", output_when="final",
                         integral_weights=[(1, 0.3)], collect_stats=True, **TOL)
    assert abs(float(res.values[0]) - known["value"]) <= 1e-13
    st = res.solver_stats
    assert (st["accepted_steps"], st["rejected_steps"], st["rhs_calls"]) == (10, 0, 61)
    assert st["solver"] == "dopri5" and st["function_evaluations"] == 61
    np.testing.assert_array_equal(res.t, [0.5])


def test_solve_lindblad_default_trace_matches_oracle():
    """Default output: the packed state at t0 and at every accepted step."""
    require_gpu()
    prepared, rho0 = _two_level(0.6)
    res = solve_lindblad(prepared, rho0, (0.0, 0.5), **TOL)
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 0.5), output="full",
                                       output_when="saveat", **TOL)
    n = int(ref["n_times"][0])
    assert res.packed_y.shape == (n, prepared.layout.packed_len) == (11, 4)
    # the embedded error estimate is a difference of O(1) stage sums, so the
    # step sizes (and the step times) agree to ~1e-8 relative, not to rounding
    np.testing.assert_allclose(res.t, ref["times"][0, :n], rtol=1e-7, atol=0)
    np.testing.assert_allclose(res.packed_y, ref["values"][0, :n], rtol=0, atol=1e-8)
    np.testing.assert_allclose(res.populations().sum(axis=1), 1.0, atol=1e-12)
    assert res.density_matrices().shape == (n, 2, 2)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_saveat_dense_output_matches_oracle(solver):
    require_gpu()
    prepared, rho0 = _two_level(0.75)
    saveat = np.linspace(0.0, 4.0, 41)
    res = solve_lindblad(prepared, rho0, (0.0, 4.0), solver=solver, saveat=saveat, output="populations", **TOL)
    ref = Oracle(prepared).solve_batch(prepared.layout.pack(rho0), (0.0, 4.0), solver=solver, saveat=saveat,
                                       output="populations", output_when="saveat", **TOL)
    np.testing.assert_allclose(res.t, saveat)
    assert res.values.shape == (41, 2)
    np.testing.assert_allclose(res.values, ref["values"][0], rtol=0, atol=1e-10)


def test_grid_scan_integral_golden_columns(golden_dir):
    """The 101-detuning study: in-solver integral, the trapezoid of 401 dense-output
    points, and the last entry of the in-solver integral trace."""
    require_gpu()
    prepared, rho0 = _two_level(0.75)
    rows = list(csv.DictReader((golden_dir / "two_level_integral_scan_101.csv").open()))
    deltas = np.linspace(-2.0, 2.0, 101)
    final = grid_scan(prepared, rho0, (0.0, 4.0), scan={"delta": deltas}, output="photon_integral",
                      integral_weights=[(1, 0.3)], **TOL)
    assert final.values.shape == (101, 1)
    assert final.metadata["grid_shape"] == (101,) and final.metadata["compact_grid"] is True
    np.testing.assert_allclose(final.values[:, 0], [float(r["in_solver_final"]) for r in rows], rtol=0, atol=1e-11)
    saveat = np.linspace(0.0, 4.0, 401)
    pops = grid_scan(prepared, rho0, (0.0, 4.0), scan={"delta": deltas}, output="populations",
                     output_when="saveat", saveat=saveat, **TOL)
    assert pops.values.shape == (101, 401, 2)
    post = np.trapezoid(0.3 * pops.values[:, :, 1], saveat, axis=1)
    np.testing.assert_allclose(post, [float(r["post_solve_trapezoid"]) for r in rows], rtol=0, atol=1e-11)
    trace = grid_scan(prepared, rho0, (0.0, 4.0), scan={"delta": deltas}, output="photon_integral",
                      integral_weights=[(1, 0.3)], output_when="saveat", saveat=saveat, **TOL)
    assert trace.values.shape == (101, 401, 1)
    np.testing.assert_allclose(trace.values[:, -1, 0], [float(r["in_solver_trace_final"]) for r in rows],
                               rtol=0, atol=1e-11)
    rate = grid_scan(prepared, rho0, (0.0, 4.0), scan={"delta": deltas}, output="photon_rate",
                     integral_weights=[(1, 0.3)], output_when="saveat", saveat=saveat, **TOL)
    np.testing.assert_allclose(rate.values[:, :, 0], 0.3 * pops.values[:, :, 1], rtol=0, atol=1e-15)


def test_batch_drivers_match_oracle():
    require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    oracle = Oracle(prepared)
    kw = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
    pb = np.asarray([[0.0, GAMMA], [1.5 * GAMMA, 0.4 * GAMMA], [-2.0 * GAMMA, 3.0 * GAMMA]], dtype=complex)
    res = parameter_scan(prepared, rho0, t_span, parameter_slots=["delta0", "Omega0"], parameter_batch=pb,
                         collect_stats=True, **kw)
    ref = oracle.solve_batch(prepared.layout.pack(rho0), t_span, parameter_slots=["delta0", "Omega0"],
                             parameter_batch=pb, output="populations", **kw)
    assert res.values.shape == (3, 39) and res.trajectory_count == 3
    np.testing.assert_allclose(res.values, ref["final"], rtol=0, atol=1e-7)
    assert res.solver_stats["accepted_steps"] == int(ref["accepted"].sum())
    assert res.parameter_slots == ["delta0", "Omega0"]
    # different initial states, selected coherences (complex), same steps as the oracle
    n = system.n_states
    rho_b = np.zeros((2, n, n), dtype=np.complex128)
    rho_b[0] = rho0
    exc = system.indices(electronic="B")
    rho_b[1, 0, 0] = 0.5
    rho_b[1, exc[0], exc[0]] = 0.5
    sel = [(0, exc[0]), (exc[0], 0), (1, 1)]
    res = initial_condition_scan(prepared, rho_b, t_span, output="selected", output_indices=sel, **kw)
    packed = prepared.layout.pack_batch(rho_b)
    ref0 = oracle.solve_batch(packed[0], t_span, output="selected", output_indices=sel, **kw)
    ref1 = oracle.solve_batch(packed[1], t_span, output="selected", output_indices=sel, **kw)
    assert res.values.dtype == np.complex128 and res.values.shape == (2, 3)
    np.testing.assert_allclose(res.values[0], ref0["final"][0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(res.values[1], ref1["final"][0], rtol=0, atol=1e-7)
    # the plain per-trajectory mode gives the same bits as the batched kernel
    plain = solve_lindblad_batch(prepared, np.stack([packed[0], packed[0], packed[0]]), t_span,
                                 parameter_slots=["delta0", "Omega0"], parameter_batch=pb,
                                 execution_mode="gpu_plain", **kw)
    batched = solve_lindblad_batch(prepared, np.stack([packed[0], packed[0], packed[0]]), t_span,
                                   parameter_slots=["delta0", "Omega0"], parameter_batch=pb, **kw)
    np.testing.assert_array_equal(plain.values, batched.values)


def test_trajectory_error_message():
    require_gpu()
    prepared, rho0 = _two_level(0.6)
    packed = prepared.layout.pack(rho0)
    with pytest.raises(ValueError, match=r"^trajectory 0: Stopped at x=[0-9.]+\. Need more than 2 steps\.$"):
        solve_lindblad_batch(prepared, np.stack([packed, packed]), (0.0, 0.5), maxiters=1, **TOL)
    with pytest.raises(ValueError, match=r"^Stopped at x=[0-9.]+\. Need more than 2 steps\.$"):
        solve_lindblad(prepared, rho0, (0.0, 0.5), maxiters=1, **TOL)
    ref = Oracle(prepared).solve_batch(packed, (0.0, 0.5), maxiters=1, raise_on_error=False, **TOL)
    x = float(re.search(r"x=(.+?)\. Need", ref["error"]).group(1))
    try:
        solve_lindblad(prepared, rho0, (0.0, 0.5), maxiters=1, **TOL)
    except ValueError as e:
        assert float(re.search(r"x=(.+?)\. Need", str(e)).group(1)) == pytest.approx(x, rel=1e-15)
