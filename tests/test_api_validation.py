"""Host-side checks of the public API: argument validation with the reference's
messages (centrex_tlf/lindblad/batch.py:170-233, solve.py:530-605), option
translation and error-message formatting.  No GPU is touched: every call here
fails before the device plan is used."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import (
    grid_scan,
    parameter_scan,
    prepare_lindblad_problem,
    solve_lindblad,
    solve_lindblad_batch,
    two_level_system,
)
from centrex_tlf_amd.lindblad import _driver as D


@pytest.fixture(scope="module")
def prepared():
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    return prepare_lindblad_problem(system, {"Omega": 0.6, "delta": 0.0}, backend="host")


@pytest.fixture(scope="module")
def rho0():
    r = np.zeros((2, 2), dtype=np.complex128)
    r[0, 0] = 1.0
    return r


def test_rust_float_formatting():
    # Rust's `{}` for f64: shortest round-trip digits without an exponent
    assert D.rust_float(1.0) == "1"
    assert D.rust_float(0.5) == "0.5"
    assert D.rust_float(1e-7) == "0.0000001"
    assert D.rust_float(2.5e-16) == "0.00000000000000025"
    assert D.rust_float(1e21) == "1000000000000000000000"
    assert D.rust_float(-0.125) == "-0.125"
    assert D.rust_float(0.1 + 0.2) == "0.30000000000000004"


def test_trajectory_messages():
    assert D.trajectory_message(1, 0.25, maxiters=1, n_plan=0) == "Stopped at x=0.25. Need more than 2 steps."
    assert D.trajectory_message(2, 3.0, maxiters=9, n_plan=0) == "Stopped at x=3. Step size underflow."
    assert D.trajectory_message(3, 1e-7, maxiters=9, n_plan=0) == "Stiff at x=0.0000001."
    assert D.trajectory_message(4, 0.0, maxiters=9, n_plan=0) == "only forward integration"
    assert D.trajectory_message(6, 3.0, maxiters=9, n_plan=7) == "ended before all saveat: 3 of 7"


def test_save_plan_matches_reference_rules():
    # t0 first when saving the start, values close to t0 dropped, range checked
    plan = D.build_save_plan(np.array([0.0, 0.1, 0.2]), 0.0, 1.0, True)
    np.testing.assert_array_equal(plan, [0.0, 0.1, 0.2])
    plan = D.build_save_plan(np.array([0.1, 0.2]), 0.0, 1.0, True)
    np.testing.assert_array_equal(plan, [0.0, 0.1, 0.2])
    plan = D.build_save_plan(np.array([0.1, 0.2]), 0.0, 1.0, False)
    np.testing.assert_array_equal(plan, [0.1, 0.2])
    assert D.build_save_plan(None, 0.0, 1.0, True) is None
    with pytest.raises(ValueError, match=r"saveat value 2 is outside \[0, 1\]"):
        D.build_save_plan(np.array([2.0]), 0.0, 1.0, False)
    with pytest.raises(ValueError, match="non-finite"):
        D.build_save_plan(np.array([np.nan]), 0.0, 1.0, False)
    # saveat step form (batch.py:47-63)
    np.testing.assert_allclose(D.normalize_saveat(0.25, (0.0, 1.0), True), [0.0, 0.25, 0.5, 0.75, 1.0])
    np.testing.assert_allclose(D.normalize_saveat(0.25, (0.0, 1.0), False), [0.25, 0.5, 0.75, 1.0])


def test_execution_modes_and_solvers(prepared, rho0):
    with pytest.raises(NotImplementedError, match="supported execution_mode values are 'reference'"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), execution_mode="nonsense")
    for mode in ("reference", "structured", "structured_upper", "expanded_sparse"):
        assert D.check_mode_and_solver(mode, "dopri5") == "persistent"
    for solver in ("scipy_bdf", "python_rk45", "rk99"):
        with pytest.raises(NotImplementedError):
            solve_lindblad(prepared, rho0, (0.0, 1.0), solver=solver)
    with pytest.raises(NotImplementedError):
        solve_lindblad(prepared, rho0, (0.0, 1.0), backend="python")
    with pytest.raises(TypeError, match="stop_event must be a RuntimeExpression"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), stop_event=object())


def test_output_validation_single(prepared, rho0):
    with pytest.raises(ValueError, match="output='selected' requires output_indices"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="selected")
    with pytest.raises(ValueError, match="output_indices is only valid with output='selected'"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="populations", output_indices=[(0, 1)])
    with pytest.raises(ValueError, match="output='photon_integral' requires integral_weights"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="photon_integral")
    with pytest.raises(ValueError, match="integral_weights are only valid with weighted output modes"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), integral_weights=[(1, 0.3)])
    with pytest.raises(ValueError, match="requires output_when='saveat'"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="photon_rate", integral_weights=[(1, 0.3)],
                       output_when="final")
    with pytest.raises(ValueError, match="requires explicit saveat values"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="photon_rate", integral_weights=[(1, 0.3)])
    with pytest.raises(NotImplementedError):
        solve_lindblad(prepared, rho0, (0.0, 1.0), output="bogus")
    with pytest.raises(ValueError, match="exactly two values"):
        solve_lindblad(prepared, rho0, (0.0, 1.0, 2.0))
    with pytest.raises(ValueError, match="dense_output=False cannot be used with interior saveat points"):
        solve_lindblad(prepared, rho0, (0.0, 1.0), saveat=[0.0, 0.5, 1.0], dense_output=False)
    with pytest.raises(TypeError, match="parameters are required"):
        solve_lindblad(two_level_system(), rho0, (0.0, 1.0))


def test_batch_validation(prepared, rho0):
    packed = prepared.layout.pack(rho0)
    batch = np.stack([packed, packed])
    with pytest.raises(ValueError, match="batch output_when='saveat' requires explicit saveat values"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), output_when="saveat")
    with pytest.raises(ValueError, match="output_when must be 'final' or 'saveat'"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), output_when="never")
    with pytest.raises(NotImplementedError):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), output="full")
    with pytest.raises(ValueError, match="parameter_slots were provided without parameter_batch"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), parameter_slots=["delta"])
    with pytest.raises(ValueError, match="parameter_slots are required"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), parameter_batch=np.zeros((2, 1)))
    with pytest.raises(ValueError, match="parameter_batch must have shape"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), parameter_slots=["delta"],
                             parameter_batch=np.zeros((3, 1)))
    with pytest.raises(ValueError, match="unknown parameter slot 'nope'"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), parameter_slots=["nope"],
                             parameter_batch=np.zeros((2, 1)))
    with pytest.raises(ValueError, match="2D rho0_batch must have shape"):
        solve_lindblad_batch(prepared, np.zeros((2, 3)), (0.0, 1.0))
    with pytest.raises(ValueError, match="threads must be positive"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), threads=0)
    with pytest.raises(ValueError, match="stop_event is only supported with output_when='final'"):
        solve_lindblad_batch(prepared, batch, (0.0, 1.0), output_when="saveat", saveat=[0.5],
                             stop_event=object())
    with pytest.raises(ValueError, match="parameter_batch must be 2D"):
        parameter_scan(prepared, rho0, (0.0, 1.0), parameter_slots=["delta"], parameter_batch=np.zeros(3))


def test_grid_scan_validation(prepared, rho0):
    with pytest.raises(ValueError, match="scan must contain at least one parameter"):
        grid_scan(prepared, rho0, (0.0, 1.0), scan={})
    with pytest.raises(ValueError, match="scan axes must be non-empty"):
        grid_scan(prepared, rho0, (0.0, 1.0), scan={"delta": []})
    with pytest.raises(TypeError, match="unexpected grid_scan keyword argument"):
        grid_scan(prepared, rho0, (0.0, 1.0), scan={"delta": [0.0]}, bogus=1)
    with pytest.raises(ValueError, match="grid_scan output_when='saveat' requires explicit saveat values"):
        grid_scan(prepared, rho0, (0.0, 1.0), scan={"delta": [0.0]}, output_when="saveat")
    with pytest.raises(ValueError, match="rho0 must be a packed vector or density matrix"):
        grid_scan(prepared, np.zeros((1, 1, 1)), (0.0, 1.0), scan={"delta": [0.0]})


def test_stop_event_normalization(prepared):
    """``events.py:27-94``: the population and expression specs and their messages."""
    import sympy as smp

    from centrex_tlf_amd.lindblad import Parameter, RuntimeExpression, normalize_stop_event, population

    with pytest.raises(ValueError, match="population event requires at least one index"):
        population([])
    with pytest.raises(ValueError, match="population event index 2 out of bounds for 2 states"):
        normalize_stop_event(population([0, 2]), prepared)
    spec = normalize_stop_event(population([1, 0, 1], threshold=0.25, name="bright"), prepared)
    assert spec == {"kind": "population", "indices": [1, 0, 1], "threshold": 0.25, "name": "bright"}
    assert normalize_stop_event(None, prepared) is None
    with pytest.raises(TypeError, match="stop_event must be a RuntimeExpression"):
        normalize_stop_event(3.0, prepared)
    t = smp.Symbol("t", real=True)
    with pytest.raises(ValueError, match=r"stop_event references unknown parameter\(s\): zeta"):
        normalize_stop_event(t - smp.Symbol("zeta"), prepared)
    spec = normalize_stop_event(RuntimeExpression(t - 0.4 * smp.Symbol("Omega")), prepared)
    assert spec["kind"] == "runtime_expression" and spec["expression"].uses_time
    slot = prepared.slot_names.index("Omega")
    assert slot in spec["expression"].slot_reads()
    # a Parameter is its own expression
    spec = normalize_stop_event(Parameter("delta"), prepared)
    assert spec["expression"].slot_reads() == {prepared.slot_names.index("delta")}


def test_entrywise_plans_and_the_expanded_sparse_check():
    """``plan_static.py:37-56``: expanded_sparse modes need a decomposed plan;
    the structured modes take an entrywise one."""
    system = two_level_system(decay=0.3, omega_name="Omega", delta_name="delta")
    entry = prepare_lindblad_problem(system, {"Omega": 0.6, "delta": 0.2}, backend="host",
                                     hamiltonian_representation="entrywise")
    assert entry.hamiltonian_plan["kind"] == "entrywise"
    with pytest.raises(ValueError, match="requires a decomposed Hamiltonian plan, but this problem was "
                                         "prepared with kind='entrywise'"):
        entry.check_execution_mode("expanded_sparse")
    entry.check_execution_mode("structured")
    entry.check_execution_mode("structured_upper")
    with pytest.raises(ValueError, match="representation must be one of"):
        prepare_lindblad_problem(system, {"Omega": 0.6}, backend="host", hamiltonian_representation="dense")
