"""The discrete adjoint on the GPU: gradients of photon counts with respect to
detuning, Rabi rate, polarisation angle and pulse width on the n=39 system
with x/y polarisation.

The adjoint differentiates the discrete solution with the accepted step sizes
held fixed.  Its exact check is therefore against central differences of the
same discrete map: a replay of the recorded step sizes (``replay``) at
perturbed parameters.  Against differences of the adaptive solve (GPU forward
and CPU oracle) the two derivatives differ by the step-size controller's own
sensitivity, which shrinks with the tolerance; those checks run at rtol 1e-12.

Errors are compared as logarithmic sensitivities p_m dI/dp_m, relative to the
largest photon count of the batch, so that a parameter whose derivative
vanishes (theta for a z-symmetric initial state) is held to the same absolute
standard as the others.  The initial state is a coherent superposition of the
ground states, which breaks the symmetry and makes the theta derivative
non-zero.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount, photon_count
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39_xy, tlf_n72_theta

pytestmark = pytest.mark.gpu

NAMES = ["delta0", "Omega0", "theta", "sigma"]
TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
TIGHT = dict(abstol=1e-14, reltol=1e-12, dt=1e-8)


def _batch(n, seed=20261202 + 4):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(-3, 3, n) * GAMMA, rng.uniform(0.3, 3, n) * GAMMA, rng.uniform(0.2, 2.9, n),
                     rng.uniform(2, 6, n) / GAMMA], axis=1)


def _coherent_ground(system, seed=7):
    """rho0 = |psi><psi| with psi a random superposition of the X(J=0) states."""
    rng = np.random.default_rng(seed)
    ground = system.indices(electronic="X", J=0)
    psi = np.zeros(system.n_states, dtype=np.complex128)
    psi[ground] = rng.normal(size=len(ground)) + 1j * rng.normal(size=len(ground))
    psi /= np.linalg.norm(psi)
    return np.outer(psi, psi.conj())


@pytest.fixture(scope="module")
def setup():
    torch = require_gpu()
    system, params, rho0, t_span, weights = tlf_n39_xy()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    y0 = prepared.layout.pack(_coherent_ground(system))
    return torch, prepared, y0, t_span, weights


def _engine(prepared, t_span, weights, tol):
    return DifferentiablePhotonCount(prepared, NAMES, weights, t_span, **tol)


def _fd(fun, p, rel=1e-6):
    out = np.zeros_like(p)
    for m in range(p.shape[1]):
        h = rel * np.abs(p[:, m])
        hi, lo = p.copy(), p.copy()
        hi[:, m] += h
        lo[:, m] -= h
        out[:, m] = (fun(hi) - fun(lo)) / (2 * h)
    return out


def _close(p, counts, got, want, rtol):
    """|p (got - want)| <= rtol * max photon count, per parameter."""
    scale = float(np.abs(counts).max())
    for m in range(want.shape[1]):
        err = np.abs(p[:, m] * (got[:, m] - want[:, m])).max()
        print(f"{NAMES[m]}: max |p dI/dp| {np.abs(p[:, m] * want[:, m]).max():.3e}  err/I {err / scale:.2e}")
        assert err <= rtol * scale, (NAMES[m], err / scale)


def test_adjoint_matches_replay_finite_differences(setup):
    torch, prepared, y0, t_span, weights = setup
    engine = _engine(prepared, t_span, weights, TOL)
    p = _batch(6)
    counts, tape = engine.forward(p, y0)
    grad = engine.gradient(p, y0, tape).cpu().numpy()
    assert np.all(np.isfinite(grad))
    # the replay at the recorded parameters is the recorded solve, bit for bit
    np.testing.assert_array_equal(engine.replay(p, y0, tape).cpu().numpy(), counts.cpu().numpy())
    fd = _fd(lambda q: engine.replay(q, y0, tape).cpu().numpy(), p)
    assert np.abs(p[:, 2] * fd[:, 2]).max() > 1e-3 * np.abs(counts.cpu().numpy()).max()  # theta matters
    _close(p, counts.cpu().numpy(), grad, fd, 1e-8)
    # the recorded forward is the plain solve's photon count
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index(n) for n in NAMES]
    ref, _, _ = plan.solve(y0, *t_span, output_kind="integral", weights=weights, override_slots=slots,
                           override_values=p.astype(complex), **TOL)
    np.testing.assert_allclose(counts.cpu().numpy(), ref.cpu().numpy()[:, 0], rtol=1e-14, atol=0)


def test_adjoint_kernel_forms_agree(setup):
    """The single-wave adjoint (variant 4, the default for this plan) and the
    multi-wave form (variant 0) compute the same derivative."""
    torch, prepared, y0, t_span, weights = setup
    engine = _engine(prepared, t_span, weights, TOL)
    p = _batch(5, seed=9)
    counts, tape = engine.forward(p, y0)
    wave = engine.gradient(p, y0, tape, variant=5).cpu().numpy()
    multi = engine.gradient(p, y0, tape, variant=1).cpu().numpy()
    auto = engine.gradient(p, y0, tape).cpu().numpy()
    np.testing.assert_array_equal(auto, wave)
    _close(p, counts.cpu().numpy(), wave, multi, 1e-12)


def test_adjoint_matches_adaptive_finite_differences(setup):
    torch, prepared, y0, t_span, weights = setup
    engine = _engine(prepared, t_span, weights, TIGHT)
    p = _batch(4, seed=5)
    counts, tape = engine.forward(p, y0)
    grad = engine.gradient(p, y0, tape).cpu().numpy()
    fd = _fd(lambda q: engine.forward(q, y0)[0].cpu().numpy(), p)
    _close(p, counts.cpu().numpy(), grad, fd, 2e-5)


def test_adjoint_matches_oracle_finite_differences(setup):
    torch, prepared, y0, t_span, weights = setup
    engine = _engine(prepared, t_span, weights, TIGHT)
    p = _batch(2, seed=11)
    counts, tape = engine.forward(p, y0)
    grad = engine.gradient(p, y0, tape).cpu().numpy()
    oracle = Oracle(prepared)

    def counts_cpu(q):
        r = oracle.solve_batch(y0, t_span, output="photon_integral", output_when="final",
                               integral_weights=weights, parameter_slots=NAMES,
                               parameter_batch=q.astype(complex), **TIGHT)
        return r["final"][:, 0]

    _close(p, counts.cpu().numpy(), grad, _fd(counts_cpu, p), 2e-5)


def test_autograd_function(setup):
    torch, prepared, y0, t_span, weights = setup
    engine = _engine(prepared, t_span, weights, TOL)
    p = torch.tensor(_batch(4, seed=3), dtype=torch.float64, device="cuda", requires_grad=True)
    target = torch.linspace(0.5, 2.0, 4, dtype=torch.float64, device="cuda")
    counts = photon_count(engine, p, y0)
    loss = ((counts - target) ** 2).sum()
    loss.backward()
    _, tape = engine.forward(p.detach(), y0)
    dIdp = engine.gradient(p.detach(), y0, tape)
    want = 2.0 * (counts.detach() - target)[:, None] * dIdp
    np.testing.assert_allclose(p.grad.cpu().numpy(), want.cpu().numpy(), rtol=1e-13, atol=0)


def test_adjoint_n72_matches_replay_finite_differences():
    """The multi-wave adjoint on the n=72 system (the 1 MHz polarisation square
    wave, theta rotating the switched pair, a coherent ground-state start)
    against central differences of the replayed discrete map."""
    require_gpu()
    system, params, rho0, _, weights = tlf_n72_theta()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    names = ["delta0", "Omega0", "theta"]
    engine = DifferentiablePhotonCount(prepared, names, weights, (0.0, 2.0 / GAMMA), tape_capacity=4096, **TOL)
    y0 = prepared.layout.pack(_coherent_ground(system))  # breaks the symmetry that zeroes d/dtheta
    p = np.asarray([[0.4 * GAMMA, 1.2 * GAMMA, 0.3], [-0.8 * GAMMA, 2.0 * GAMMA, 1.1]])
    counts, tape = engine.forward(p, y0)
    grad = engine.gradient(p, y0, tape).cpu().numpy()
    assert np.all(np.isfinite(grad))
    fd = _fd(lambda q: engine.replay(q, y0, tape).cpu().numpy(), p)
    c = counts.cpu().numpy()
    scale = float(np.abs(c).max())
    for m, name in enumerate(names):
        err = np.abs(p[:, m] * (grad[:, m] - fd[:, m])).max()
        print(f"{name}: max |p dI/dp| {np.abs(p[:, m] * fd[:, m]).max():.3e}  err/I {err / scale:.2e}")
        assert err <= 1e-7 * scale, (name, err / scale)
    assert np.abs(p[:, 2] * fd[:, 2]).max() > 1e-4 * scale  # theta matters here
