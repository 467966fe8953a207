"""Solves and adjoints of one plan from two threads on two streams at once.

The reference's batch and grid calls release the GIL
(rust/src/lindblad/python_api.rs:619, 748), so concurrent calls on one
prepared problem are legal there.  Here every call takes its own work-queue
counter, and adjoints of one plan are ordered on the device around their
shared workspace, so overlapping calls give bitwise the single-stream results.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.adjoint import DifferentiablePhotonCount

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39, tlf_n39_xy

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)


def _run_threads(work):
    errors = []

    def guarded(fn):
        try:
            fn()
        except Exception as exc:  # surfaced in the main thread
            errors.append(exc)

    threads = [threading.Thread(target=guarded, args=(fn,)) for fn in work]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads)
    if errors:
        raise errors[0]


def test_two_streams_solve_one_plan():
    torch = require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    plan = prepared.device_plan()
    slots = [prepared.slot_names.index("delta0"), prepared.slot_names.index("Omega0")]
    rng = np.random.default_rng(31)
    batches = {name: np.stack([rng.uniform(-5, 5, n) * GAMMA, rng.uniform(0.1, 10, n) * GAMMA], axis=1).astype(complex)
               for name, n in (("a", 384), ("b", 257))}
    y0 = torch.as_tensor(prepared.layout.pack(rho0), device="cuda")
    kw = dict(output_kind="populations", weights=weights, with_integral=True, override_slots=slots, **TOL)
    want = {k: [t.cpu() for t in plan.solve(y0, *t_span, override_values=v, **kw)] for k, v in batches.items()}
    got: dict[str, list] = {}

    def worker(name):
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            outs = [plan.solve(y0, *t_span, override_values=batches[name], stream=stream, **kw) for _ in range(6)]
        stream.synchronize()
        got[name] = [[t.cpu() for t in o] for o in outs]

    _run_threads([lambda: worker("a"), lambda: worker("b")])
    for name in batches:
        for rep, outs in enumerate(got[name]):
            for g, w in zip(outs, want[name]):
                assert torch.equal(g.nan_to_num(7.0), w.nan_to_num(7.0)), (name, rep)


def test_two_streams_adjoint_one_plan():
    torch = require_gpu()
    system, params, rho0, t_span, weights = tlf_n39_xy()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    engine = DifferentiablePhotonCount(prepared, ["delta0", "Omega0", "theta", "sigma"], weights, t_span, **TOL)
    y0 = prepared.layout.pack(rho0)
    rng = np.random.default_rng(32)
    batches = {name: np.stack([rng.uniform(-3, 3, n) * GAMMA, rng.uniform(0.3, 3, n) * GAMMA,
                               rng.uniform(0.2, 2.9, n), rng.uniform(2, 6, n) / GAMMA], axis=1)
               for name, n in (("a", 96), ("b", 64))}
    tapes = {k: engine.forward(v, y0)[1] for k, v in batches.items()}
    want = {k: engine.gradient(v, y0, tapes[k]).cpu() for k, v in batches.items()}
    got: dict[str, list] = {}

    def worker(name):
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            outs = [engine.gradient(batches[name], y0, tapes[name]) for _ in range(4)]
        stream.synchronize()
        got[name] = [o.cpu() for o in outs]

    _run_threads([lambda: worker("a"), lambda: worker("b")])
    for name in batches:
        for rep, g in enumerate(got[name]):
            assert torch.equal(g, want[name]), (name, rep)
