"""Plan-specialised paired kernels (``lb_plan_specialize``): the time-dependent
coefficient programs compiled into the kernel at run time give bitwise the
results of the expression machine.

Cases: the n=39 TlF grid workload (Gaussian envelope: time, constant power,
real exp) under Dopri5 and Tsit5, and a two-level system whose coefficients
use a complex parameter, a builtin (sin), helpers (phase_modulation, the
square-wave profile), a complex power with a non-constant exponent and
constant real powers.  A plan without time dependence is left on the
precompiled kernels."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import (
    LindbladParameters,
    Time,
    gaussian,
    phase_modulation,
    prepare_lindblad_problem,
    sine,
    square_wave_profile,
    two_level_system,
)
from centrex_tlf_amd.lindblad.native import DevicePlan, GridAxes
from oracle.lb_oracle import Oracle

from ._gpu import require_gpu
from .systems import GAMMA, tlf_n39

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    for x, y, name in zip(a, b, ("values", "stats", "fail_t", "integral")):
        assert x.shape == y.shape, (what, name)
        assert bool((x.nan_to_num(7.0) == y.nan_to_num(7.0)).all()), (what, name)


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_n39_grid_is_bitwise_the_expression_machine(solver):
    torch = require_gpu()
    system, params, rho0, t_span, weights = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    base = DevicePlan(prepared.flat_plan)
    spec = DevicePlan(prepared.flat_plan).specialize()
    y0_np = prepared.layout.pack(rho0)
    y0 = torch.as_tensor(y0_np, device="cuda")
    names = prepared.slot_names
    axes = GridAxes([np.linspace(-5.0, 5.0, 24) * GAMMA, np.geomspace(0.1, 10.0, 24) * GAMMA])
    kw = dict(solver=solver, output_kind="populations", weights=weights, with_integral=True,
              override_slots=[names.index("delta0"), names.index("Omega0")], grid_axes=axes,
              support=base.support_of(y0_np), abstol=1e-10, reltol=1e-8, dt=1e-8)
    want = base.solve(y0, *t_span, **kw)
    got = spec.solve(y0, *t_span, **kw)
    got2 = spec.solve(y0, *t_span, **kw)  # the compiled kernel is reused
    torch.cuda.synchronize()
    assert spec.specialized
    _same(got, want, "n39")
    _same(got2, want, "n39 again")
    assert int(want[1][:, 0].min()) > 10  # the trajectories took many steps


def _rich_two_level():
    p = LindbladParameters()
    a = p.complex("a", 0.6 + 0.25j)
    w = p.real("w", 1.3)
    t = Time()
    p.bind("Omega", a * sine(t, angular_frequency=w) + phase_modulation(t, 0.5, w) / (1.0 + t) ** w
           + square_wave_profile(t, low=0.1, high=0.4, period=2.0), finalize=False)
    p.bind("delta", -((w * t + 0.01) ** 0.5) * 0.1 + gaussian(t, center=2.0, sigma=0.7, amplitude=0.3),
           finalize=False)
    p._finalize()
    return two_level_system(omega_name="Omega", delta_name="delta"), p


@pytest.mark.parametrize("solver", ["dopri5", "tsit5"])
def test_rich_expressions_are_bitwise_the_expression_machine(solver):
    torch = require_gpu()
    system, params = _rich_two_level()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    base = DevicePlan(prepared.flat_plan)
    spec = DevicePlan(prepared.flat_plan).specialize()
    rho0 = np.diag([1.0, 0.0]).astype(complex)
    y0 = torch.as_tensor(prepared.layout.pack(rho0), device="cuda")
    names = prepared.slot_names
    rng = np.random.default_rng(20261202 + 21)
    ov = np.stack([rng.uniform(0.2, 1.5, 37) + 1j * rng.uniform(-0.5, 0.5, 37),
                   rng.uniform(0.5, 2.5, 37) + 0j], axis=1)
    kw = dict(solver=solver, output_kind="full", weights=[(1, 0.3)], with_integral=True,
              override_slots=[names.index("a"), names.index("w")], override_values=ov,
              abstol=1e-10, reltol=1e-8, dt=1e-3)
    want = base.solve(y0, 0.0, 6.0, **kw)
    got = spec.solve(y0, 0.0, 6.0, **kw)
    torch.cuda.synchronize()
    _same(got, want, "two-level")
    assert np.all(np.isfinite(got[0].cpu().numpy()))
    # and the oracle on the same expressions.  The right-hand side, where the
    # expressions enter, agrees to rounding at 40 times across the span; the
    # solves agree to 1e-6: the square-wave drive jumps, and steps that land
    # an ulp apart (the device's sin/exp round apart from the host's) cross
    # the jumps with different local errors
    oracle = Oracle(prepared)
    ts = np.linspace(0.013, 5.97, 40)
    ys = np.repeat(prepared.layout.pack(np.array([[0.6, 0.2 + 0.1j], [0.2 - 0.1j, 0.4]]))[None], ts.size, 0)
    rhs = base.rhs(torch.as_tensor(ys, device="cuda"), torch.as_tensor(ts, device="cuda")).cpu().numpy()
    for b in range(ts.size):
        want = oracle.rhs(ys[b], ts[b])
        np.testing.assert_allclose(rhs[b], want, rtol=0, atol=1e-12 * np.abs(want).max(), err_msg=f"t={ts[b]}")
    ref = oracle.solve_batch(prepared.layout.pack(rho0), (0.0, 6.0), solver=solver, output="full",
                             output_when="final", parameter_slots=["a", "w"], parameter_batch=ov,
                             abstol=1e-10, reltol=1e-8, dt=1e-3)
    np.testing.assert_allclose(got[0].cpu().numpy(), ref["final"], rtol=0, atol=1e-6)


def test_static_plan_is_left_alone():
    torch = require_gpu()
    system = two_level_system(omega_name="Omega", delta_name="delta")
    prepared = prepare_lindblad_problem(system, {"Omega": 0.7, "delta": 0.2}, backend="gpu")
    plan = DevicePlan(prepared.flat_plan).specialize()
    y0 = torch.as_tensor(prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex)), device="cuda")
    ref = DevicePlan(prepared.flat_plan).solve(y0, 0.0, 5.0, output_kind="populations", abstol=1e-10,
                                                reltol=1e-8, dt=1e-3)
    got = plan.solve(y0, 0.0, 5.0, output_kind="populations", abstol=1e-10, reltol=1e-8, dt=1e-3)
    torch.cuda.synchronize()
    _same(got, ref, "static")


def test_first_call_inside_a_capture_takes_the_precompiled_kernel():
    """No compilation inside a stream capture: the captured first call runs the
    precompiled kernel, a later eager call compiles; every result is bitwise
    the expression machine's."""
    torch = require_gpu()
    system, params = _rich_two_level()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    base = DevicePlan(prepared.flat_plan)
    spec = DevicePlan(prepared.flat_plan).specialize()
    dev = torch.device("cuda")
    y0_np = prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex))
    y0 = torch.as_tensor(y0_np, device=dev)
    names = prepared.slot_names
    ov = torch.as_tensor(np.stack([np.linspace(0.3, 1.2, 16) + 0.1j, np.linspace(0.6, 2.0, 16) + 0j], axis=1),
                         device=dev)
    kw = dict(output_kind="populations", weights=[(1, 0.3)], with_integral=True,
              override_slots=[names.index("a"), names.index("w")], override_values=ov, workgroups=8,
              support=base.support_of(y0_np), abstol=1e-10, reltol=1e-8, dt=1e-3)  # no host copy in the capture
    want = [t.clone() for t in base.solve(y0, 0.0, 6.0, **kw)]
    spec.specialize(False)  # builds the plan's restricted tables (allocations) before the capture
    spec.solve(y0, 0.0, 6.0, **kw)
    spec.specialize(True)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        captured = spec.solve(y0, 0.0, 6.0, stream=stream, **kw)
    torch.cuda.synchronize()
    for _ in range(3):
        for t in captured:
            t.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        _same(captured, want, "captured first call")
    _same(spec.solve(y0, 0.0, 6.0, **kw), want, "eager call after the capture")


def test_failed_compilation_runs_the_precompiled_kernel(monkeypatch):
    """A run-time compilation that fails (forced here through the compile
    options hook: a macro redefinition made an error) leaves the solve on the
    precompiled kernel with bitwise the same results, warns once with the
    reason, and is not attempted again for the plan."""
    torch = require_gpu()
    system, params = _rich_two_level()
    prepared = prepare_lindblad_problem(system, params, backend="gpu")
    base = DevicePlan(prepared.flat_plan)
    monkeypatch.setenv("CENTREX_LB_RTC_OPTIONS", "-Werror -DLB_PART=1")
    spec = DevicePlan(prepared.flat_plan).specialize()
    y0 = torch.as_tensor(prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex)), device="cuda")
    names = prepared.slot_names
    ov = np.stack([np.linspace(0.3, 1.2, 24) + 0.1j, np.linspace(0.6, 2.0, 24) + 0j], axis=1)
    kw = dict(output_kind="populations", weights=[(1, 0.3)], with_integral=True,
              override_slots=[names.index("a"), names.index("w")], override_values=ov,
              abstol=1e-10, reltol=1e-8, dt=1e-3)
    want = base.solve(y0, 0.0, 6.0, **kw)
    assert spec.specialize_failure() is None
    with pytest.warns(UserWarning, match="precompiled kernel used"):
        got = spec.solve(y0, 0.0, 6.0, **kw)
    torch.cuda.synchronize()
    _same(got, want, "failed compilation")
    why = spec.specialize_failure()
    assert why is not None and "lb_solve_pair_kernel" in why, why
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("error")  # neither a second compile nor a second warning
        again = spec.solve(y0, 0.0, 6.0, **kw)
    torch.cuda.synchronize()
    _same(again, want, "after the failure")
