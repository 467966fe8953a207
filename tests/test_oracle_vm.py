"""Pins of the oracle's expression machine (rust/src/lindblad/eval.rs:135-451,
664-911) by closed forms that do not restate its formulas.

Every scalar helper id and every builtin is evaluated by the oracle at many
times and compared with:

- the Julia definitions the reference's own tests transcribe for the
  waveforms (tests/lindblad/test_parameter_helper_wrappers.py:138-160,
  226-240): square_wave, sawtooth_wave, variable_on_off_duty;
- library routines (numpy's complex sin/cos/tan/exp/abs/power, scipy's normal
  density, scipy.constants for the physical constants);
- identities a slip of sign, index or operand breaks: the rotated Gaussian in
  its principal axes, 0.5 (cos + sin) = sin(x + pi/4) / sqrt 2, the parity of
  floor(u) as the sign of sin(pi u), the on/off gate as a phase of the period.

Points within 1e-9 of a waveform's switching instant are skipped, as the
reference's tests skip them.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.lb_oracle import eval_expression

from . import _closed_forms as cf
from ._closed_forms import TIMES

CONST, SLOT, TIME, ADD, SUB, MUL, DIV, POW, NEG, CONJ, BUILTIN, HELPER = 1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13


def helper(fid: int, t: float, *args: float) -> complex:
    prog = [(TIME, 0, 0, 0, 0.0)] + [(CONST, 0, 0, 0, a) for a in args] + [(HELPER, 0, 1 + len(args), fid, 0.0)]
    return eval_expression(prog, t)


def helper_x(fid: int, *args: float) -> complex:
    prog = [(CONST, 0, 0, 0, a) for a in args] + [(HELPER, 0, len(args), fid, 0.0)]
    return eval_expression(prog)


def test_square_wave_is_the_julia_definition():
    checked = 0
    for t in TIMES:
        want = cf.julia_square_wave(t, 1.2, 0.3)
        if want is None:
            continue
        assert helper(4, t, 1.2, 0.3) == want, t
        checked += 1
    assert checked > 0.9 * len(TIMES)


def test_sawtooth_wave_is_the_julia_definition():
    for t in TIMES:
        want = cf.julia_sawtooth_wave(t, 1.2, 0.3)
        if want is None:
            continue
        got = helper(6, t, 1.2, 0.3)
        assert got.imag == 0.0
        assert got.real == pytest.approx(want, abs=1e-12), t


@pytest.mark.parametrize("duty,inv_period,phase", [(0.25, 0.4, 0.0), (0.5, 1.3, 0.7), (0.9, 0.31, -1.1)])
def test_variable_on_off_duty_is_the_julia_definition(duty, inv_period, phase):
    for t in TIMES:
        want = cf.julia_on_off_duty(t, duty, inv_period, phase)
        if want is not None:
            assert helper(12, t, duty, inv_period, phase) == want, t


@pytest.mark.parametrize("ton,toff,phase", [(0.3, 0.7, 0.0), (1.1, 0.4, 2.0), (0.05, 0.5, -0.4)])
def test_variable_on_off_is_on_for_ton_of_each_period(ton, toff, phase):
    on_time = 0.0
    for t in TIMES:
        got = helper(7, t, ton, toff, phase)
        want = cf.on_off(t, ton, toff, phase)
        if want is not None:
            assert got == want, t
        on_time += got.real
    assert on_time / len(TIMES) == pytest.approx(ton / (ton + toff), abs=0.05)


def test_alternating_sign_is_the_sign_of_sin_pi_u():
    for t in TIMES:
        want = cf.alternating_sign(t, 0.4, 0.9)
        if want is not None:
            assert helper(13, t, 0.4, 0.9) == want, t


def test_gaussian_1d_is_the_normal_density():
    for t in TIMES[::5]:
        assert helper(15, t, 0.3, 0.7).real == pytest.approx(cf.gaussian_1d(t, 0.3, 0.7), rel=1e-12, abs=1e-300)


def test_gaussian_2d_factorises():
    rng = np.random.default_rng(5)
    for _ in range(50):
        x, y, mx, my = rng.normal(size=4)
        a, sx, sy = rng.uniform(0.2, 2.0, size=3)
        assert helper_x(1, x, y, a, mx, my, sx, sy).real == pytest.approx(cf.gaussian_2d(x, y, a, mx, my, sx, sy), rel=1e-12)


def test_gaussian_2d_rotated_in_its_principal_axes():
    rng = np.random.default_rng(6)
    for _ in range(50):
        x, y, mx, my = rng.normal(size=4)
        a, sx, sy = rng.uniform(0.2, 2.0, size=3)
        th = rng.uniform(-math.pi, math.pi)
        want = cf.gaussian_2d_rotated(x, y, a, mx, my, sx, sy, th)
        assert helper_x(2, x, y, a, mx, my, sx, sy, th).real == pytest.approx(want, rel=1e-11)
    # theta = 0 is the axis-aligned Gaussian; a point on the rotated x axis sees sigma_x only
    assert helper_x(2, 0.4, -0.2, 1.5, 0.1, 0.3, 0.6, 1.1, 0.0) == helper_x(1, 0.4, -0.2, 1.5, 0.1, 0.3, 0.6, 1.1)
    r, th = 0.8, 0.7
    got = helper_x(2, 0.1 + r * math.cos(th), 0.3 + r * math.sin(th), 1.0, 0.1, 0.3, 0.6, 1.1, th).real
    assert got == pytest.approx(math.exp(-r * r / (2 * 0.6 ** 2)), rel=1e-12)


def test_phase_modulation_is_a_unit_phasor():
    beta, omega = 0.8, 2.1
    for t in TIMES[::3]:
        got = helper(3, t, beta, omega)
        assert abs(got - cf.phase_modulation(t, beta, omega)) <= 1e-15


def test_resonant_polarization_modulation_identity():
    gamma, omega = 1.3, 0.9
    for t in TIMES[::3]:
        want = cf.resonant_polarization_modulation(t, gamma, omega)
        assert abs(helper(5, t, gamma, omega) - want) <= 2e-15


def test_rabi_from_intensity_uses_the_codata_constants():
    for intensity, coupling in [(1.0, 0.5), (250.0, -0.3), (3.2e4, 1.0)]:
        # CODATA 2018 vs 2022 epsilon_0 differ by 7e-10 relative
        want = cf.rabi_from_intensity(intensity, coupling, 2.0e-30)
        assert helper_x(9, intensity, coupling, 2.0e-30).real == pytest.approx(want, rel=1e-8)
        assert helper_x(9, intensity, coupling).real == pytest.approx(cf.rabi_from_intensity(intensity, coupling), rel=1e-8)


def test_gaussian_beam_rabi_is_the_composition():
    args = (0.2, -0.1, 120.0, 0.0, 0.05, 0.4, 0.6)
    inten = helper_x(1, *args).real
    assert helper_x(11, *args, 0.7).real == pytest.approx(helper_x(9, inten, 0.7).real, rel=1e-14)
    assert helper_x(11, *args, 0.7, 3e-30).real == pytest.approx(helper_x(9, inten, 0.7, 3e-30).real, rel=1e-14)


@pytest.mark.parametrize("fid,fn", [(1, np.sin), (2, np.cos), (3, np.tan), (4, np.exp)])
def test_complex_builtins_are_the_library_functions(fid, fn):
    rng = np.random.default_rng(fid)
    for z in rng.normal(size=20) + 1j * rng.normal(size=20):
        got = eval_expression([(CONST, 0, 0, 0, z), (BUILTIN, 0, 1, fid, 0.0)])
        want = complex(fn(z))
        assert abs(got - want) <= 2e-15 * max(1.0, abs(want)), z  # a few ulp: num-complex vs numpy formulas


def test_abs_real_imag_conj_and_arithmetic():
    z, w = 0.3 - 1.7j, -2.2 + 0.4j
    ev = lambda *prog: eval_expression(list(prog))
    c = lambda v: (CONST, 0, 0, 0, v)
    assert ev(c(z), (BUILTIN, 0, 1, 5, 0.0)) == pytest.approx(abs(z), rel=1e-16)
    assert ev(c(z), (BUILTIN, 0, 1, 6, 0.0)) == z.real
    assert ev(c(z), (BUILTIN, 0, 1, 7, 0.0)) == z.imag
    assert ev(c(z), (CONJ, 0, 0, 0, 0.0)) == z.conjugate()
    assert ev(c(z), (NEG, 0, 0, 0, 0.0)) == -z
    assert ev(c(z), c(w), (SUB, 0, 0, 0, 0.0)) == z - w  # operand order: left - right
    assert ev(c(z), c(w), (DIV, 0, 0, 0, 0.0)) == pytest.approx(z / w, rel=1e-15)
    for base, ex in [(z, w), (2.5, 0.5), (-2.0, 0.5), (-3.0, 2.0), (z, 3.0)]:
        got = ev(c(base), c(ex), (POW, 0, 0, 0, 0.0))
        assert abs(got - complex(base) ** complex(ex)) <= 1e-14 * abs(complex(base) ** complex(ex)), (base, ex)
    assert ev(c(0.0), c(0.0), (POW, 0, 0, 0, 0.0)) == 1.0  # x^0 = 1
    assert ev(c(0.0), c(2.0), (POW, 0, 0, 0, 0.0)) == 0.0
    # slots and time
    assert eval_expression([(SLOT, 1, 0, 0, 0.0), (TIME, 0, 0, 0, 0.0), (MUL, 0, 0, 0, 0.0)], 2.0, [0j, 1.5 - 1j]) == 3.0 - 2j
    # comparisons give 1/0 on the real parts
    for op, f in [(14, lambda a, b: a > b), (15, lambda a, b: a >= b), (16, lambda a, b: a < b),
                  (17, lambda a, b: a <= b), (18, lambda a, b: a == b), (19, lambda a, b: a != b)]:
        for a, b in [(1.0, 2.0), (2.0, 2.0), (3.0, 2.0)]:
            assert ev(c(a), c(b), (op, 0, 0, 0, 0.0)) == (1.0 if f(a, b) else 0.0)
