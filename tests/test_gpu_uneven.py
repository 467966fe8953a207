"""Uneven batches on the GPU: the n=39 system and the n=72 system (two lasers
with x/y polarisation switched by a 1 MHz square wave, microwave J=1-J=2) in
one call, each trajectory with its own pulse width sigma, centre sigma and end
time 2 sigma (sigma ~ N(5, 1.5)/Gamma clipped above 1/Gamma), Dopri5 and Tsit5
at rtol 1e-8.

Checks: the grouped call reproduces each group solved alone bit for bit; the
autotuned launch reproduces the default launch bit for bit (all variants
round identically); a seeded subset matches the CPU oracle (populations and
photon counts to 1e-7 absolute, accepted-step counts equal); a HIP graph of
the grouped launch replays it exactly; failures name group and trajectory.
"""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import UnevenGroup, prepare_lindblad_problem, solve_uneven_batch
from centrex_tlf_amd.lindblad.native import solve_group
from centrex_tlf_amd.lindblad.uneven import _group_calls, autotune
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39, tlf_n72

pytestmark = pytest.mark.gpu

TOL = dict(abstol=1e-10, reltol=1e-8, dt=1e-8)
NAMES = ["sigma", "t_center", "delta0"]


def _rows(count, seed):
    rng = np.random.default_rng(seed)
    sigma = np.clip(rng.normal(5.0, 1.5, count), 1.0, None) / GAMMA
    delta = rng.uniform(-2.0, 2.0, count) * GAMMA
    return np.stack([sigma, sigma, delta], axis=1).astype(complex), 2.0 * sigma


@pytest.fixture(scope="module")
def systems():
    require_gpu()
    out = []
    for build in (tlf_n39, tlf_n72):
        system, params, rho0, _, weights = build()
        out.append((prepare_lindblad_problem(system, params, backend="gpu"), rho0, weights))
    return out


def _groups(systems, counts, seed=20261202 + 3):
    groups = []
    for k, ((prepared, rho0, weights), count) in enumerate(zip(systems, counts)):
        pb, t_end = _rows(count, seed + k)
        groups.append(UnevenGroup(prepared, rho0, t_end, NAMES, pb, weights))
    return groups


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_grouped_call_reproduces_groups_alone(systems, solver):
    groups = _groups(systems, (40, 12))
    res = solve_uneven_batch(groups, solver=solver, tuning=None, **TOL)
    for grp, r in zip(groups, res):
        plan = grp.prepared.device_plan()
        y0 = grp.prepared.layout.pack(grp.rho0)
        slots = [grp.prepared.slot_names.index(n) for n in NAMES]
        pops, stats, _, phot = plan.solve(y0, 0.0, grp.t_end, solver=solver, output_kind="populations",
                                          weights=grp.integral_weights, with_integral=True,
                                          override_slots=slots, override_values=grp.parameter_batch, **TOL)
        np.testing.assert_array_equal(r.populations, pops.cpu().numpy())
        np.testing.assert_array_equal(r.photons, phot.cpu().numpy())
        np.testing.assert_array_equal(r.stats, stats.cpu().numpy())
        assert np.all(r.stats[:, 3] == 0)
        assert r.stats[:, 0].max() > 1.5 * r.stats[:, 0].min()  # uneven step counts


def test_autotuned_launch_is_bitwise_default(systems):
    groups = _groups(systems, (64, 8), seed=5)
    tun = autotune(groups, repeats=1, cache=False, **TOL)
    assert len(tun.variants) == 2 and all(v > 0 for v in tun.variants)
    assert tun.mode in ("sequential", "concurrent") and np.isfinite(tun.measured_ms)
    print("tuning:", tun.mode, tun.variants, tun.threads, tun.workgroups, f"{tun.measured_ms:.3f} ms")
    tuned = solve_uneven_batch(groups, tuning=tun, **TOL)
    plain = solve_uneven_batch(groups, tuning=None, longest_first=False, **TOL)
    for a, b in zip(tuned, plain):
        np.testing.assert_array_equal(a.populations, b.populations)
        np.testing.assert_array_equal(a.photons, b.photons)
        np.testing.assert_array_equal(a.stats, b.stats)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_uneven_matches_oracle(systems, solver):
    groups = _groups(systems, (6, 3), seed=11)
    res = solve_uneven_batch(groups, solver=solver, tuning=None, **TOL)
    for grp, r in zip(groups, res):
        oracle = Oracle(grp.prepared)
        y0 = grp.prepared.layout.pack(grp.rho0)
        # n=39: first and last; n=72 (about 6 s of oracle time each): the shortest
        picks = (0, len(grp.t_end) - 1) if grp.prepared.layout.n < 64 else (int(np.argmin(grp.t_end)),)
        for b in picks:
            span = (0.0, float(grp.t_end[b]))
            kw = dict(solver=solver, output_when="final", parameter_slots=NAMES,
                      parameter_batch=grp.parameter_batch[b:b + 1], **TOL)
            ref = oracle.solve_batch(y0, span, output="populations", **kw)
            ref_i = oracle.solve_batch(y0, span, output="photon_integral",
                                       integral_weights=grp.integral_weights, **kw)
            np.testing.assert_allclose(r.populations[b], ref["final"][0], rtol=0, atol=1e-7)
            np.testing.assert_allclose(r.photons[b], ref_i["final"][0, 0], rtol=0, atol=1e-7)
            assert int(r.stats[b, 0]) == int(ref["accepted"][0])
            assert abs(r.populations[b].sum() - 1.0) <= 1e-10


def test_grouped_graph_replays(systems):
    torch = require_gpu()
    groups = _groups(systems, (32, 8), seed=13)
    calls = _group_calls(groups, t0=0.0, solver="dopri5", maxiters=100_000, save_start=True,
                         longest_first=True, **TOL)
    dev = torch.device("cuda")
    for _, kw in calls:  # device inputs, so the captured graph copies nothing from the host
        kw["y0"] = torch.as_tensor(kw["y0"], device=dev)
        kw["t1"] = torch.as_tensor(kw["t1"], device=dev)
        kw["override_values"] = torch.as_tensor(kw["override_values"], device=dev)
        kw["queue_order"] = torch.as_tensor(kw["queue_order"], device=dev)
        kw["workgroups"] = 64
    eager = solve_group(calls)
    torch.cuda.synchronize()
    eager = [[t.clone() for t in out] for out in eager]
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        solve_group(calls, stream=stream)  # warm-up on the capture stream
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        captured = solve_group(calls, stream=stream)
    torch.cuda.synchronize()
    for rep in range(3):
        for out in captured:
            for t in out:
                t.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for got, want in zip(captured, eager):
            for g, w in zip(got, want):
                assert torch.equal(g, w), f"replay {rep}"


def test_failure_names_group_and_trajectory(systems):
    groups = _groups(systems, (4, 2), seed=17)
    with pytest.raises(ValueError, match=r"group 0, trajectory \d+: Stopped at x=.*Need more than 6 steps"):
        solve_uneven_batch(groups, tuning=None, abstol=1e-10, reltol=1e-8, dt=1e-8, maxiters=5)


def test_every_multi_wave_variant_is_bitwise_default(systems):
    """Each multi-wave solver variant that covers the restricted n=72 plan
    (slots per lane 1 to 12, one to sixteen waves) gives bitwise the default
    launch's results on short transits.  The variants the autotuner may pick
    are exactly these, so a variant that rounds or orders differently fails
    here rather than inside a tuned batch."""
    from centrex_tlf_amd.lindblad.native import kernel_variants

    prepared, rho0, weights = systems[1]
    plan = prepared.device_plan()
    y0 = prepared.layout.pack(rho0)
    pb, _ = _rows(6, 11)
    slots = [prepared.slot_names.index(n) for n in NAMES]
    support = plan.support_of(y0)
    kw = dict(output_kind="populations", weights=weights, with_integral=True, override_slots=slots,
              override_values=pb, support=support, **TOL)
    t_end = 0.6 / GAMMA
    want = [t.cpu().numpy() for t in plan.solve(y0, 0.0, t_end, **kw)]
    S = plan.support_info(support)["slot_count"]
    checked = []
    for v in kernel_variants():  # the single-wave and paired forms cover 64 slots and are skipped
        if v["slots_per_thread"] * v["max_threads"] < S:
            continue
        print(f"variant {v}", flush=True)
        got = [t.cpu().numpy() for t in plan.solve(y0, 0.0, t_end, variant=v["variant"], **kw)]
        for a, b in zip(got, want):
            np.testing.assert_array_equal(np.nan_to_num(a, nan=7.0), np.nan_to_num(b, nan=7.0), err_msg=str(v))
        checked.append(v["variant"])
    assert len(checked) >= 8, checked
