"""The device expression machine against closed forms (the GPU mirror of
tests/test_oracle_vm.py).

Each case binds the two-level Rabi frequency to one helper or builtin of time,
H = [[0, Omega/2], [Omega*/2, 0]], and evaluates the device RHS kernel at the
state |0><0|: there d rho_01/dt = i Omega(t) / 2 exactly (a product with 1/2,
i and 1.0 and sums with exact zeros), so the kernel returns the coefficient
the device machine computed.  That value is compared, at the reference
tests' 401 times, with the closed forms of tests/_closed_forms.py: the Julia
waveform definitions, library routines and identities
(rust/src/lindblad/eval.rs:135-451 is what is being checked).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import sympy as smp

from centrex_tlf_amd.lindblad import (
    LindbladParameters,
    Time,
    alternating_sign,
    gaussian_1d,
    gaussian_2d,
    gaussian_2d_rotated,
    gaussian_beam_rabi,
    phase_modulation,
    prepare_lindblad_problem,
    resonant_polarization_modulation,
    sawtooth_wave,
    square_wave,
    two_level_system,
    variable_on_off,
    variable_on_off_duty,
)
from centrex_tlf_amd.lindblad.parameters import RuntimeExpression
from oracle.lb_oracle import Oracle

from . import _closed_forms as cf
from ._closed_forms import TIMES
from ._gpu import require_gpu

pytestmark = pytest.mark.gpu


def _builtin(fn):
    t = Time()
    return RuntimeExpression(fn(t.expr), {})


CASES = [
    # name, expression of Time(), closed form, exact (waveforms) or relative tolerance
    ("square_wave", lambda: square_wave(Time(), 1.2, 0.3), lambda t: cf.julia_square_wave(t, 1.2, 0.3), 0.0),
    ("sawtooth_wave", lambda: sawtooth_wave(Time(), 1.2, 0.3), lambda t: cf.julia_sawtooth_wave(t, 1.2, 0.3), 1e-12),
    ("on_off_duty_a", lambda: variable_on_off_duty(Time(), 0.25, 0.4, 0.0), lambda t: cf.julia_on_off_duty(t, 0.25, 0.4, 0.0), 0.0),
    ("on_off_duty_b", lambda: variable_on_off_duty(Time(), 0.9, 0.31, -1.1), lambda t: cf.julia_on_off_duty(t, 0.9, 0.31, -1.1), 0.0),
    ("on_off", lambda: variable_on_off(Time(), 1.1, 0.4, 2.0), lambda t: cf.on_off(t, 1.1, 0.4, 2.0), 0.0),
    ("alternating_sign", lambda: alternating_sign(Time(), 0.4, 0.9), lambda t: cf.alternating_sign(t, 0.4, 0.9), 0.0),
    ("gaussian_1d", lambda: gaussian_1d(Time(), 0.3, 7.0), lambda t: cf.gaussian_1d(t, 0.3, 7.0), 1e-12),
    ("gaussian_2d", lambda: gaussian_2d(Time(), 0.2, 1.5, 0.1, 0.3, 6.0, 1.1),
     lambda t: cf.gaussian_2d(t, 0.2, 1.5, 0.1, 0.3, 6.0, 1.1), 1e-12),
    ("gaussian_2d_rotated", lambda: gaussian_2d_rotated(Time(), 0.2, 1.5, 0.1, 0.3, 6.0, 1.1, 0.7),
     lambda t: cf.gaussian_2d_rotated(t, 0.2, 1.5, 0.1, 0.3, 6.0, 1.1, 0.7), 1e-11),
    ("phase_modulation", lambda: phase_modulation(Time(), 0.8, 2.1), lambda t: cf.phase_modulation(t, 0.8, 2.1), 1e-15),
    ("resonant_polarization", lambda: resonant_polarization_modulation(Time(), 1.3, 0.9),
     lambda t: cf.resonant_polarization_modulation(t, 1.3, 0.9), 2e-15),
    ("gaussian_beam_rabi", lambda: gaussian_beam_rabi(Time(), 0.2, 120.0, 0.0, 0.05, 9.0, 0.6, 0.7),
     lambda t: cf.rabi_from_intensity(cf.gaussian_2d(t, 0.2, 120.0, 0.0, 0.05, 9.0, 0.6), 0.7), 1e-8),
    ("sin_cos_exp", lambda: _builtin(lambda t: smp.sin(0.3 * t) * smp.exp(smp.I * 0.2 * t) + smp.cos(t)),
     lambda t: complex(np.sin(0.3 * t) * np.exp(1j * 0.2 * t) + np.cos(t)), 1e-14),
    ("tan_abs", lambda: _builtin(lambda t: smp.tan(0.05 * t + 0.1j) + smp.Abs(t - 3)),
     lambda t: complex(np.tan(0.05 * t + 0.1j) + abs(t - 3)), 1e-14),
]


@pytest.mark.parametrize("name,make,closed,tol", CASES, ids=[c[0] for c in CASES])
def test_device_expression_machine_is_the_closed_form(name, make, closed, tol):
    torch = require_gpu()
    p = LindbladParameters()
    p.bind("Omega", make(), finalize=False)
    p.bind("delta", 0.0, finalize=False)
    p._finalize()
    prepared = prepare_lindblad_problem(two_level_system(omega_name="Omega", delta_name="delta"), p, backend="host")
    plan = prepared.device_plan()
    y = prepared.layout.pack(np.diag([1.0, 0.0]).astype(complex))
    ts = np.asarray(TIMES)
    out = plan.rhs(torch.as_tensor(np.tile(y, (len(ts), 1)), device="cuda"),
                   torch.as_tensor(ts, device="cuda")).cpu().numpy()
    omega = 2.0 * out[:, 3] - 2.0j * out[:, 2]
    oracle = Oracle(prepared)
    checked = 0
    for i, t in enumerate(ts):
        want = closed(float(t))
        if want is None:
            continue
        want = complex(want)
        if tol == 0.0:
            assert omega[i] == want, (name, t)
        else:
            assert abs(omega[i] - want) <= tol * max(abs(want), 1e-300) + 1e-300, (name, t, omega[i], want)
        checked += 1
        # and the oracle's machine on the same input, to rounding
        ref = oracle.rhs(y, float(t))
        assert abs(complex(2.0 * ref[3], -2.0 * ref[2]) - omega[i]) <= 4e-15 * max(abs(want), 1e-300), (name, t)
    assert checked > 0.9 * len(ts)
