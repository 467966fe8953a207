"""The n = 72 uneven-batch workload against the oracle on a seeded sample.

n = 72: X(J=0..3) + B(J=1, F1'=3/2, F'=1,2), two lasers whose polarisation
switches x <-> y with a 1 MHz square wave, a microwave J=1-J=2 at 0.5 Gamma.
64 trajectories drawn as the uneven batch draws them: sigma ~ N(5, 1.5)/Gamma
clipped above 1/Gamma, envelope centre sigma, end time 2 sigma, detuning
~ U[-2, 2] Gamma; Dopri5 and Tsit5 at rtol 1e-8, atol 1e-10.

Bars: final populations and photon counts within 1e-7 absolute of the oracle
(run as one batch on all host threads, populations and the photon trapezoid
from one pass), |Tr rho - 1| <= 1e-10, and equal accepted-step counts.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import UnevenGroup, prepare_lindblad_problem, solve_uneven_batch
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n72

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
NAMES = ["sigma", "t_center", "delta0"]


@pytest.fixture(scope="module")
def sample():
    require_gpu()
    system, params, rho0, _, weights = tlf_n72()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    rng = np.random.default_rng(20261202 + 2)
    sigma = np.clip(rng.normal(5.0, 1.5, 512), 1.0, None) / GAMMA
    delta = rng.uniform(-2.0, 2.0, 512) * GAMMA
    take = np.sort(rng.choice(512, size=64, replace=False))
    pb = np.stack([sigma, sigma, delta], axis=1).astype(complex)[take]
    return UnevenGroup(prepared, rho0, 2.0 * sigma[take], NAMES, pb, weights)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_n72_sample_matches_oracle(sample, solver):
    (res,) = solve_uneven_batch([sample], solver=solver, tuning=None, **TOL)
    oracle = Oracle(sample.prepared)
    y0 = sample.prepared.layout.pack(sample.rho0)
    ref = oracle.solve_batch(y0, (0.0, float(sample.t_end.max())), solver=solver, output="populations",
                             output_when="final", integral_weights=sample.integral_weights, with_integral=True,
                             parameter_slots=NAMES, parameter_batch=sample.parameter_batch,
                             t1_per_traj=sample.t_end, threads=0, **TOL)
    assert np.all(ref["status"] == 0) and np.all(res.stats[:, 3] == 0)
    pop_err = np.abs(res.populations - ref["final"]).max(axis=1)
    ph_err = np.abs(res.photons - ref["integral"])
    trace_err = np.abs(res.populations.sum(axis=1) - 1.0)
    print(f"{solver}: 64 n=72 trajectories, accepted {int(res.stats[:, 0].min())}..{int(res.stats[:, 0].max())}; "
          f"max |population error| {pop_err.max():.2e}, max |photon error| {ph_err.max():.2e}, "
          f"max |trace - 1| {trace_err.max():.2e}")
    np.testing.assert_array_equal(res.stats[:, 0], ref["accepted"])
    assert pop_err.max() <= 1e-7
    assert ph_err.max() <= 1e-7
    assert trace_err.max() <= 1e-10
