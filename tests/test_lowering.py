"""Host lowering + oracle RHS against the textbook dense Lindblad formula.

The dense side substitutes parameter values into the sympy Hamiltonian with
lambdify, so it shares nothing with the bytecode compiler, the decomposed
lowering or the sparse program.  Systems: the reference's two-level and chain
fixtures (chain n=3 and n=45 straddle the 40-state kernel gate) and the two
TlF systems of the workloads.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import PackedHermitianLayout, prepare_lindblad_problem, two_level_system
from centrex_tlf_amd.lindblad.ir import evaluate_parameter_graph
from oracle.lb_oracle import Oracle

from ._dense import CHAIN_PARAMETERS, chain_system, dense_rhs, random_density
from .systems import tlf_n39, tlf_n72


def test_packed_layout_roundtrip():
    layout = PackedHermitianLayout(3)
    rho = np.array([[0.2, 0.1 + 0.3j, -0.4j], [0.1 - 0.3j, 0.5, 0.2 - 0.1j], [0.4j, 0.2 + 0.1j, 0.3]])
    packed = layout.pack(rho)
    np.testing.assert_allclose(layout.unpack(packed), rho)
    assert packed[layout.upper_real_index(0, 2)] == 0.0
    assert packed[layout.upper_imag_index(0, 2)] == -0.4
    assert packed[layout.upper_imag_index(1, 2)] == -0.1


def _systems():
    two = two_level_system(omega_name="Omega", delta_name="delta")
    yield "two_level", two, {"Omega": 0.6, "delta": 0.3}, 0.0
    yield "chain3", chain_system(3), dict(CHAIN_PARAMETERS), 0.41
    yield "chain45", chain_system(45), dict(CHAIN_PARAMETERS), 0.41


@pytest.mark.parametrize("name,system,params,t", list(_systems()), ids=lambda v: v if isinstance(v, str) else "")
def test_rhs_matches_dense(name, system, params, t):
    prepared = prepare_lindblad_problem(system, params, backend="host")
    oracle = Oracle(prepared)
    rho = random_density(system.n_states, seed=0)
    got = oracle.rhs(prepared.layout.pack(rho), t)
    want = prepared.layout.pack(dense_rhs(prepared, rho, t))
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=1e-9)


@pytest.mark.parametrize("builder", [tlf_n39, tlf_n72], ids=["tlf_n39", "tlf_n72"])
def test_tlf_rhs_matches_dense(builder):
    system, params, rho0, t_span, _ = builder()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    oracle = Oracle(prepared)
    rho = random_density(system.n_states, seed=3) * 1e-2
    t = 0.37 * t_span[1]
    got = oracle.rhs(prepared.layout.pack(rho), t)
    want = prepared.layout.pack(dense_rhs(prepared, rho, t))
    # magnitude-aware tolerance: GHz-scale diagonal energies leave float64 residue ~eps*|H|*|rho|
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, atol=64 * np.finfo(float).eps * scale, rtol=1e-9)
    # a density matrix keeps unit trace: the diagonal of the RHS sums to zero
    d = oracle.rhs(prepared.layout.pack(rho0), t)
    assert abs(d[: system.n_states].sum()) <= 1e-6 * np.abs(d).max() + 1e-300


def test_parameter_graph_time_dependence():
    system, params, _, t_span, _ = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    graph = prepared.parameter_graph
    slots = evaluate_parameter_graph(graph, 10.0 / (2 * np.pi * 1.56e6))
    omega_slot = graph["slot_names"].index("Ω0")
    assert abs(slots[omega_slot] - 2 * np.pi * 1.56e6) < 1e-6
    coeffs = prepared.hamiltonian_plan["coefficients"]
    assert len(coeffs) == 2
    assert any(c["expression"].uses_time for c in coeffs) is False  # time enters via the Ω0 compound
    assert any(e.uses_time for _, e in graph["compounds"])


def test_term_counts_and_dedup():
    system, params, _, _, _ = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    rhs = prepared.expanded_rhs_plan
    n = system.n_states
    assert rhs["output_ptrs"].size == n * (n + 1) // 2 + 1
    assert rhs["output_ptrs"][-1] == rhs["input_indices"].size
    assert np.all(np.hypot(rhs["factors_re"], rhs["factors_im"]) > 1e-15)
    # keys are unique per output
    for o in range(n * (n + 1) // 2):
        a, b = rhs["output_ptrs"][o], rhs["output_ptrs"][o + 1]
        keys = list(zip(rhs["input_indices"][a:b], rhs["coeff_indices"][a:b], rhs["coeff_conj"][a:b],
                        rhs["input_conj"][a:b]))
        assert len(keys) == len(set(keys)) and keys == sorted(keys)


@pytest.mark.parametrize("name,system,params,t", list(_systems())[:2], ids=["two_level", "chain3"])
def test_entrywise_plan_rhs_matches_dense(name, system, params, t):
    """The entrywise lowering (one coefficient per distinct upper entry) runs
    the same equation as the dense formula and as the decomposed plan."""
    prepared = prepare_lindblad_problem(system, params, backend="host", hamiltonian_representation="entrywise")
    decomposed = prepare_lindblad_problem(system, params, backend="host")
    rho = random_density(system.n_states, seed=1)
    got = Oracle(prepared).rhs(prepared.layout.pack(rho), t)
    want = prepared.layout.pack(dense_rhs(prepared, rho, t))
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=1e-9)
    np.testing.assert_allclose(got, Oracle(decomposed).rhs(prepared.layout.pack(rho), t), atol=1e-14, rtol=1e-12)


def test_entrywise_plan_rhs_matches_dense_tlf_n39():
    system, params, rho0, t_span, _ = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="host", hamiltonian_representation="entrywise")
    assert prepared.hamiltonian_plan["kind"] == "entrywise"
    rho = random_density(system.n_states, seed=5) * 1e-2
    t = 0.41 * t_span[1]
    got = Oracle(prepared).rhs(prepared.layout.pack(rho), t)
    want = prepared.layout.pack(dense_rhs(prepared, rho, t))
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, atol=1e-12 * scale, rtol=0)


def test_structured_jumps_of_single_jump_operators():
    """Rates |C|^2 of each operator's one element C[target, source]: per-jump
    records, the decay rate summed per source and the transfers per target."""
    from centrex_tlf_amd.lindblad.plan_static import lower_structured_jumps

    c = np.zeros((3, 3, 3), dtype=np.complex128)
    c[0, 0, 2] = 0.6 + 0.8j  # 2 -> 0, rate 1.0
    c[1, 1, 2] = 2.0  # 2 -> 1, rate 4.0
    c[2, 0, 1] = -0.5j  # 1 -> 0, rate 0.25
    jumps, decay, incoming = lower_structured_jumps(c)
    assert [(j["target"], j["source"]) for j in jumps] == [(0, 2), (1, 2), (0, 1)]
    np.testing.assert_allclose([j["rate"] for j in jumps], [1.0, 4.0, 0.25], rtol=1e-15)
    assert (jumps[0]["amp_re"], jumps[0]["amp_im"]) == (0.6, 0.8)
    np.testing.assert_allclose(decay, [0.0, 0.25, 5.0], rtol=1e-15)
    assert [[t["source"] for t in row] for row in incoming] == [[2, 1], [2], []]
    c[1, 0, 0] = 1.0  # a second element in operator 1
    with pytest.raises(AssertionError, match="collapse operator 1 is not a single-jump operator; found 2"):
        lower_structured_jumps(c)
