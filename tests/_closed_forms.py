"""Closed forms of the expression machine's helpers that do not restate the
reference's formulas (rust/src/lindblad/eval.rs:155-451, 664-789).

Each function returns the expected value at one time, or None for a time
within 1e-9 of a waveform's switching instant (the reference's own tests skip
those: tests/lindblad/test_parameter_helper_wrappers.py:163-181).  Shared by
the oracle pins (tests/test_oracle_vm.py) and the device pins
(tests/test_gpu_vm.py).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.constants as sc
from scipy.stats import norm

GUARD = 1e-9
TIMES = [i * 0.137 for i in range(-200, 201)]
ED_XTB = 0.315 * sc.e * sc.physical_constants["Bohr radius"][0]


def _near(u: float, period: float, edges) -> bool:
    u = u % period
    return any(min(abs(u - e), abs(u - e - period), abs(u - e + period)) < GUARD for e in edges)


def julia_square_wave(t, omega, phase):
    """julia_common.jl 0.5 (1 + squarewave(x)); Waveforms.jl squarewave(x) =
    mod2pi(x) < pi ? 1 : -1 (transcribed by the reference's tests)."""
    x = omega * t + phase
    if _near(x, 2 * math.pi, (0.0, math.pi)):
        return None
    return 1.0 if (x % (2 * math.pi)) < math.pi else 0.0


def julia_sawtooth_wave(t, omega, phase):
    """0.5 (1 + sawtoothwave(x - pi)), sawtoothwave(x) = rem2pi(x, RoundNearest) / pi."""
    if _near(omega * t + phase, 2 * math.pi, (0.0,)):
        return None
    x = omega * t + phase - math.pi
    return 0.5 * (1.0 + (x - 2 * math.pi * round(x / (2 * math.pi))) / math.pi)


def julia_on_off_duty(t, duty, inv_period, phase):
    """mod1(t invT + phase / 2pi, 1) < duty (julia_common.jl)."""
    frac = (t * inv_period + phase / (2 * math.pi)) % 1.0
    if frac <= 0.0:
        frac += 1.0
    if min(abs(frac - duty), frac, abs(1.0 - frac)) < GUARD:
        return None
    return 1.0 if frac < duty else 0.0


def on_off(t, ton, toff, phase):
    """On for the first ton of each period ton + toff, the period shifted by phase / 2pi of itself."""
    period = ton + toff
    u = (t + phase * period / (2 * math.pi)) % period
    if _near(u, period, (0.0, ton)):
        return None
    return 1.0 if u < ton else 0.0


def alternating_sign(t, x0, width):
    """(-1)^floor(u) with u = (t - x0) / width is the sign of sin(pi u)."""
    u = (t - x0) / width
    if abs(u - round(u)) < GUARD:
        return None
    return math.copysign(1.0, math.sin(math.pi * u))


def gaussian_1d(x, center, sigma):
    return norm.pdf(x, loc=center, scale=sigma) * sigma * math.sqrt(2 * math.pi)


def gaussian_2d(x, y, a, mx, my, sx, sy):
    return a * gaussian_1d(x, mx, sx) * gaussian_1d(y, my, sy)


def gaussian_2d_rotated(x, y, a, mx, my, sx, sy, th):
    """The Gaussian in its principal axes, rotated by +th."""
    dx, dy = x - mx, y - my
    u = math.cos(th) * dx + math.sin(th) * dy
    v = -math.sin(th) * dx + math.cos(th) * dy
    return a * math.exp(-(u * u / (2 * sx * sx) + v * v / (2 * sy * sy)))


def phase_modulation(t, beta, omega):
    return complex(np.exp(1j * beta * np.sin(omega * t)))


def resonant_polarization_modulation(t, gamma, omega):
    """0.5 (cos th + sin th) (1 + i) = sin(th + pi/4) (1 + i) / sqrt 2, th = gamma sin(omega t) / 2."""
    v = math.sin(0.5 * gamma * math.sin(omega * t) + math.pi / 4) / math.sqrt(2.0)
    return complex(v, v)


def rabi_from_intensity(intensity, coupling, dipole=ED_XTB):
    field = math.sqrt(2 * intensity / (sc.c * sc.epsilon_0))
    return field * coupling * dipole / sc.hbar
