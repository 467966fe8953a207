"""Execution modes and Hamiltonian representations on the GPU.

The reference's ``reference``, ``structured`` and ``structured_upper`` modes
are host algorithms for the same equation as ``expanded_sparse``
(rust/src/lindblad/rhs.rs:1280-1460); here they run the plan's expanded
sparse program, so their results are those of ``expanded_sparse``.  Entrywise
plans (one coefficient per distinct upper Hamiltonian entry,
centrex_tlf/lindblad/ir.py:616-674) are checked against the oracle running
the same plan, and against the decomposed plan (the same equation; they
differ by rounding in the coefficient arithmetic)."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem, solve_lindblad_batch
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-9)


@pytest.fixture(scope="module")
def n39_both():
    require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    dec = prepare_lindblad_problem(system, params, backend="gpu")
    ent = prepare_lindblad_problem(system, params, backend="gpu", hamiltonian_representation="entrywise")
    rng = np.random.default_rng(20261202 + 13)
    pb = np.stack([rng.uniform(-5, 5, 8) * GAMMA, rng.uniform(0.1, 10, 8) * GAMMA], axis=1).astype(complex)
    return dec, ent, dec.layout.pack(rho0), t_span, weights, pb


def test_entrywise_plan_matches_the_oracle_and_the_decomposed_plan(n39_both):
    dec, ent, y0, t_span, weights, pb = n39_both
    assert ent.hamiltonian_plan["kind"] == "entrywise"
    slots = [ent.slot_names.index("delta0"), ent.slot_names.index("Omega0")]
    kw = dict(output_kind="populations", weights=weights, with_integral=True, override_slots=slots,
              override_values=pb, **TOL)
    pops, stats, _, phot = ent.device_plan().solve(y0, *t_span, **kw)
    ref = Oracle(ent).solve_batch(y0, t_span, output="populations", output_when="final",
                                  parameter_slots=["delta0", "Omega0"], parameter_batch=pb, **TOL)
    np.testing.assert_allclose(pops.cpu().numpy(), ref["final"], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(stats.cpu().numpy()[:, 0], ref["accepted"])
    dpops, _, _, dphot = dec.device_plan().solve(y0, *t_span, **kw)
    np.testing.assert_allclose(pops.cpu().numpy(), dpops.cpu().numpy(), rtol=0, atol=1e-7)
    np.testing.assert_allclose(phot.cpu().numpy(), dphot.cpu().numpy(), rtol=1e-6)


def test_reference_mode_strings_run_the_expanded_sparse_program(n39_both):
    dec, ent, y0, t_span, weights, pb = n39_both
    batch = np.stack([y0] * pb.shape[0])
    kw = dict(parameter_slots=["delta0", "Omega0"], parameter_batch=pb, output="populations", **TOL)
    base = solve_lindblad_batch(dec, batch, t_span, execution_mode="expanded_sparse", **kw)
    for mode in ("reference", "structured", "structured_upper"):
        res = solve_lindblad_batch(dec, batch, t_span, execution_mode=mode, **kw)
        np.testing.assert_array_equal(res.values, base.values)
    # an entrywise plan takes the structured modes, and keeps the reference's
    # error for the expanded_sparse one
    res = solve_lindblad_batch(ent, batch, t_span, execution_mode="structured", **kw)
    np.testing.assert_allclose(res.values, base.values, rtol=0, atol=1e-7)
    with pytest.raises(ValueError, match="requires a decomposed Hamiltonian plan"):
        solve_lindblad_batch(ent, batch, t_span, execution_mode="expanded_sparse", **kw)
