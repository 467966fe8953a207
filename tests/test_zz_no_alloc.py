"""lb_adjoint_noalloc: the adjoint with its workspace in the library's fixed
device array, called through the C ABI, against lb_adjoint on the same tape.

Two calls in a row, the second with fewer trajectories: the array keeps the
first call's contents, so the second shows that what must start at zero is
cleared.  Each trajectory's gradient is computed by one workgroup with the
same kernel, so the results are bitwise those of lb_adjoint whatever the
number of workgroups.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39_xy, tlf_n72_theta

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)


def _batch(n, seed, cols):
    rng = np.random.default_rng(seed)
    out = np.stack([rng.uniform(-3, 3, n) * GAMMA, rng.uniform(0.3, 3, n) * GAMMA, rng.uniform(0.2, 2.9, n),
                    rng.uniform(2, 6, n) / GAMMA], axis=1)
    return out[:, :cols]


@pytest.mark.parametrize("mode", ["adaptive", "frozen"])
def test_noalloc_adjoint_matches_lb_adjoint(mode):
    torch = require_gpu()
    system, params, rho0, t_span, weights = tlf_n39_xy()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    engine = DifferentiablePhotonCount(prepared, ["delta0", "Omega0", "theta", "sigma"], weights, t_span,
                                       mode=mode, **TOL)
    y0 = prepared.layout.pack(rho0)
    for n, seed in ((700, 1), (37, 2)):
        p = _batch(n, seed, 4)
        _, tape = engine.forward(p, y0)
        want = engine.gradient(p, y0, tape)
        got = engine.gradient(p, y0, tape, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (mode, n, float((got - want).abs().max()))


def test_noalloc_adjoint_multi_wave_plan():
    """The n = 72 plan (multi-wave adjoint, about 1 MB of workspace per
    workgroup): the array holds fewer workgroups, with the same results."""
    torch = require_gpu()
    system, params, rho0, _, weights = tlf_n72_theta()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    engine = DifferentiablePhotonCount(prepared, ["delta0", "Omega0", "theta"], weights, (0.0, 1.0 / GAMMA),
                                       tape_capacity=4096, **TOL)
    y0 = prepared.layout.pack(rho0)
    for n, seed in ((96, 3), (5, 4)):
        p = _batch(n, seed, 3)
        _, tape = engine.forward(p, y0)
        want = engine.gradient(p, y0, tape)
        got = engine.gradient(p, y0, tape, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, float((got - want).abs().max()))
