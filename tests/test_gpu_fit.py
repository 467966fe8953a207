"""Batched photon-count fit on the GPU (forward with tape + discrete adjoint +
Adam each iteration) over an uneven mix of the n=39 x/y system and the n=72
system with a polarisation angle: targets come from seeded true parameters,
the fit starts from perturbed ones, and the loss must fall."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount
from centrex_tlf_amd.lindblad.fit import FitGroup, fit_photon_counts

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39_xy, tlf_n72_theta

pytestmark = pytest.mark.gpu

NAMES = ["delta0", "Omega0", "theta"]
TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)


def test_fit_reduces_the_loss():
    require_gpu()
    rng = np.random.default_rng(20261202 + 5)
    groups = []
    for build, count, t_scale in ((tlf_n39_xy, 12, 1.0), (tlf_n72_theta, 3, 0.2)):
        system, params, rho0, t_span, weights = build()
        prepared = prepare_lindblad_problem(system, params, backend="host")
        sigma = np.clip(rng.normal(5.0, 1.5, count), 1.0, None) / GAMMA * t_scale
        t_end = 2.0 * sigma
        true = np.stack([rng.uniform(-1, 1, count) * GAMMA, rng.uniform(0.5, 2, count) * GAMMA,
                         rng.uniform(0, np.pi, count)], axis=1)
        fixed = {"sigma": sigma.astype(complex), "t_center": sigma.astype(complex)}
        eng = DifferentiablePhotonCount(prepared, NAMES, weights, (0.0, t_end), fixed=fixed, tape_capacity=8192,
                                        **TOL)
        targets, _ = eng.forward(true, prepared.layout.pack(rho0))
        initial = true * np.asarray([1.0, 1.25, 1.0]) + np.asarray([0.3 * GAMMA, 0.0, 0.2])
        groups.append(FitGroup(prepared, rho0, t_end, targets.cpu().numpy(), initial, weights, fixed))
    res = fit_photon_counts(groups, NAMES, scales=[GAMMA, GAMMA, 1.0], iterations=8, lr=0.05,
                            tape_capacity=8192, **TOL)
    print("loss", [f"{v:.3e}" for v in res.loss], f"{res.gradients_per_second:.1f} gradients/s",
          f"tape {res.tape_bytes / 1e6:.1f} MB")
    assert np.all(np.isfinite(res.loss))
    assert res.loss[-1] < 0.5 * res.loss[0]
    final = sum(float(((c - g.targets) ** 2).sum()) for c, g in zip(res.counts, groups))
    assert final < res.loss[0]


def test_graph_replayed_fit_is_the_eager_fit_and_matches_the_oracle():
    """The fit with every iteration after the first replayed from one HIP
    graph (forward solves, adjoints, loss and Adam step of both systems)
    follows the eager fit, and the final populations of 10 trajectories (8
    n=39 x/y, 2 n=72) agree with the CPU oracle within 1e-7."""
    require_gpu()
    from oracle.lb_oracle import Oracle

    rng = np.random.default_rng(20261202 + 6)
    groups = []
    for build, count, t_scale in ((tlf_n39_xy, 8, 1.0), (tlf_n72_theta, 2, 0.2)):
        system, params, rho0, t_span, weights = build()
        prepared = prepare_lindblad_problem(system, params, backend="host")
        sigma = np.clip(rng.normal(5.0, 1.5, count), 1.0, None) / GAMMA * t_scale
        t_end = 2.0 * sigma
        true = np.stack([rng.uniform(-1, 1, count) * GAMMA, rng.uniform(0.5, 2, count) * GAMMA,
                         rng.uniform(0, np.pi, count)], axis=1)
        fixed = {"sigma": sigma.astype(complex), "t_center": sigma.astype(complex)}
        eng = DifferentiablePhotonCount(prepared, NAMES, weights, (0.0, t_end), fixed=fixed, tape_capacity=8192,
                                        **TOL)
        targets, _ = eng.forward(true, prepared.layout.pack(rho0))
        initial = true * np.asarray([1.0, 1.25, 1.0]) + np.asarray([0.3 * GAMMA, 0.0, 0.2])
        groups.append(FitGroup(prepared, rho0, t_end, targets.cpu().numpy(), initial, weights, fixed))
    kw = dict(scales=[GAMMA, GAMMA, 1.0], iterations=5, lr=0.05, tape_capacity=8192, **TOL)
    eager = fit_photon_counts(groups, NAMES, **kw)
    graph = fit_photon_counts(groups, NAMES, graph=True, **kw)
    print("eager loss", eager.loss, "\ngraph loss", graph.loss)
    np.testing.assert_allclose(graph.loss, eager.loss, rtol=1e-9, atol=0)
    for a, b in zip(graph.parameters, eager.parameters):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=0)
    assert graph.loss[-1] < graph.loss[0]
    for grp, p in zip(groups, graph.parameters):
        names = NAMES + ["sigma", "t_center"]
        vals = np.concatenate([p, np.stack([grp.fixed["sigma"].real, grp.fixed["t_center"].real], axis=1)], axis=1)
        dev = grp.prepared.device_plan()
        y0 = grp.prepared.layout.pack(grp.rho0)
        slots = [grp.prepared.slot_names.index(m) for m in names]
        pops = dev.solve(y0, 0.0, grp.t_end, output_kind="populations", override_slots=slots,
                         override_values=vals.astype(complex), **TOL)[0].cpu().numpy()
        ref = Oracle(grp.prepared).solve_batch(y0, (0.0, 1.0), output="populations", output_when="final",
                                               parameter_slots=names, parameter_batch=vals.astype(complex),
                                               t1_per_traj=grp.t_end, **TOL)
        np.testing.assert_allclose(pops, ref["final"], rtol=0, atol=1e-7)
