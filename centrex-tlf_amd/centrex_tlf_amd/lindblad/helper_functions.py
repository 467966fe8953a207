"""Waveform and beam helpers that runtime parameter expressions may call.

The numeric ids are part of the plan format: the device expression machine
dispatches on them, so they equal the reference's ids
(``centrex_tlf/lindblad/helper_functions.py:35-51``,
``rust/src/lindblad/eval.rs:664-789``).  The Python bodies are the closed
forms of those helpers and are used for host-side evaluation (plotting,
``RuntimeExpression.evaluate``) and by the host expression evaluator.
"""

from __future__ import annotations

import math
from enum import IntEnum
from typing import Callable, Iterable, Mapping

__all__ = [
    "ED_XTB",
    "HelperFunctionId",
    "HELPER_FUNCTION_ALIASES",
    "HELPER_FUNCTION_IDS",
    "HELPER_FUNCTION_NAMES",
    "HELPER_FUNCTIONS",
    "TUPLE_HELPER_IDS",
]

# Transition dipole moment of the X-B band, 0.315 e a0 in C m
# (centrex_tlf/constants/constants.py:386).
_ELEMENTARY_CHARGE = 1.602176634e-19
_BOHR_RADIUS = 5.29177210544e-11
ED_XTB = 0.315 * _ELEMENTARY_CHARGE * _BOHR_RADIUS


class HelperFunctionId(IntEnum):
    GAUSSIAN_2D = 1
    GAUSSIAN_2D_ROTATED = 2
    PHASE_MODULATION = 3
    SQUARE_WAVE = 4
    RESONANT_POLARIZATION_MODULATION = 5
    SAWTOOTH_WAVE = 6
    VARIABLE_ON_OFF = 7
    MULTIPASS_2D_INTENSITY = 8
    RABI_FROM_INTENSITY = 9
    MULTIPASS_2D_RABI = 10
    GAUSSIAN_BEAM_RABI = 11
    VARIABLE_ON_OFF_DUTY = 12
    ALTERNATING_SIGN = 13
    LINEAR_INTERP = 14
    GAUSSIAN_1D = 15
    PCHIP_INTERP = 16


# Helpers that take tuple arguments; the device expression machine evaluates
# scalar helpers only.
TUPLE_HELPER_IDS = frozenset({8, 10, 14, 16})


def gaussian_2d(x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y):
    dx = x - mean_x
    dy = y - mean_y
    return amplitude * math.exp(
        -((dx * dx) / (2.0 * sigma_x * sigma_x) + (dy * dy) / (2.0 * sigma_y * sigma_y))
    )


def gaussian_2d_rotated(x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y, theta):
    c = math.cos(theta)
    s = math.sin(theta)
    a = c * c / (2.0 * sigma_x * sigma_x) + s * s / (2.0 * sigma_y * sigma_y)
    b = math.sin(2.0 * theta) / (2.0 * sigma_x * sigma_x) - math.sin(2.0 * theta) / (
        2.0 * sigma_y * sigma_y
    )
    cc = s * s / (2.0 * sigma_x * sigma_x) + c * c / (2.0 * sigma_y * sigma_y)
    dx = x - mean_x
    dy = y - mean_y
    return amplitude * math.exp(-(a * dx * dx + b * dx * dy + cc * dy * dy))


def phase_modulation(t, beta, omega):
    phi = beta * math.sin(omega * t)
    return complex(math.cos(phi), math.sin(phi))


def square_wave(t, omega, phase):
    return 0.5 * (1.0 + (1.0 if math.sin(omega * t + phase) >= 0.0 else -1.0))


def resonant_polarization_modulation(t, gamma, omega):
    theta = 0.5 * gamma * math.sin(omega * t)
    a = 0.5 * (math.cos(theta) + math.sin(theta))
    return complex(a, a)


def sawtooth_wave(t, omega, phase):
    return ((omega * t + phase) / (2.0 * math.pi)) % 1.0


def variable_on_off(t, ton, toff, phase):
    period = ton + toff
    duty = ton / period
    frac = ((2.0 * math.pi * t / period + phase) % (2.0 * math.pi)) / (2.0 * math.pi)
    return 1.0 if frac < duty else 0.0


def variable_on_off_duty(t, duty, inv_period, phase):
    # Half-open (0, 1] phase fraction: an exact 0 counts as the end of a cycle.
    frac = (t * inv_period + phase / (2.0 * math.pi)) % 1.0
    if frac <= 0.0:
        frac += 1.0
    return 1.0 if frac < duty else 0.0


def multipass_2d_intensity(x, y, amplitudes, xlocs, ylocs, sigma_x, sigma_y):
    total = 0.0
    for amplitude, xloc, yloc in zip(amplitudes, xlocs, ylocs):
        total += gaussian_2d(x, y, amplitude, xloc, yloc, sigma_x, sigma_y)
    return total


def rabi_from_intensity(intensity, coupling, dipole_moment=ED_XTB):
    hbar = 1.0545718176461565e-34
    c = 299792458.0
    eps0 = 8.8541878128e-12
    electric_field = math.sqrt(intensity * 2.0 / (c * eps0))
    return electric_field * coupling * dipole_moment / hbar


def multipass_2d_rabi(x, y, intensities, xlocs, ylocs, sigma_x, sigma_y, main_coupling,
                      dipole_moment=ED_XTB):
    intensity = multipass_2d_intensity(x, y, intensities, xlocs, ylocs, sigma_x, sigma_y)
    return rabi_from_intensity(intensity, main_coupling, dipole_moment)


def gaussian_beam_rabi(x, y, intensity, xloc, yloc, sigma_x, sigma_y, main_coupling,
                       dipole_moment=ED_XTB):
    beam = gaussian_2d(x, y, intensity, xloc, yloc, sigma_x, sigma_y)
    return rabi_from_intensity(beam, main_coupling, dipole_moment)


def alternating_sign(x, x0, width):
    k = math.floor((x - x0) / width)
    return 1.0 if k % 2 == 0 else -1.0


def linear_interp(x, grid: Iterable[float], values: Iterable[float]):
    g = [float(v) for v in grid]
    v = [float(w) for w in values]
    if len(g) != len(v):
        raise ValueError("linear_interp grid and values must have the same length")
    if not g:
        raise ValueError("linear_interp grid must contain at least one point")
    if len(g) == 1 or x <= g[0]:
        return v[0]
    if any(b <= a for a, b in zip(g, g[1:])):
        raise ValueError("linear_interp grid must be strictly increasing")
    for i in range(len(g) - 1):
        if x <= g[i + 1]:
            return v[i] + (x - g[i]) / (g[i + 1] - g[i]) * (v[i + 1] - v[i])
    return v[-1]


def gaussian_1d(x, center, sigma):
    dx = x - center
    return math.exp(-(dx * dx) / (2.0 * sigma * sigma))


def pchip_interp(x, grid: Iterable[float], values: Iterable[float]):
    from scipy.interpolate import PchipInterpolator

    return float(PchipInterpolator([float(a) for a in grid], [float(b) for b in values],
                                   extrapolate=True)(x))


HELPER_FUNCTIONS: Mapping[str, Callable[..., complex | float]] = {
    "gaussian_1d": gaussian_1d,
    "gaussian_2d": gaussian_2d,
    "gaussian_2d_rotated": gaussian_2d_rotated,
    "phase_modulation": phase_modulation,
    "square_wave": square_wave,
    "resonant_polarization_modulation": resonant_polarization_modulation,
    "sawtooth_wave": sawtooth_wave,
    "variable_on_off": variable_on_off,
    "multipass_2d_intensity": multipass_2d_intensity,
    "rabi_from_intensity": rabi_from_intensity,
    "multipass_2d_rabi": multipass_2d_rabi,
    "gaussian_beam_rabi": gaussian_beam_rabi,
    "variable_on_off_duty": variable_on_off_duty,
    "variable_on_off_duty_invT": variable_on_off_duty,
    "alternating_sign": alternating_sign,
    "linear_interp": linear_interp,
    "pchip_interp": pchip_interp,
}

HELPER_FUNCTION_IDS: Mapping[str, HelperFunctionId] = {
    "gaussian_1d": HelperFunctionId.GAUSSIAN_1D,
    "gaussian_2d": HelperFunctionId.GAUSSIAN_2D,
    "gaussian_2d_rotated": HelperFunctionId.GAUSSIAN_2D_ROTATED,
    "phase_modulation": HelperFunctionId.PHASE_MODULATION,
    "square_wave": HelperFunctionId.SQUARE_WAVE,
    "resonant_polarization_modulation": HelperFunctionId.RESONANT_POLARIZATION_MODULATION,
    "sawtooth_wave": HelperFunctionId.SAWTOOTH_WAVE,
    "variable_on_off": HelperFunctionId.VARIABLE_ON_OFF,
    "multipass_2d_intensity": HelperFunctionId.MULTIPASS_2D_INTENSITY,
    "rabi_from_intensity": HelperFunctionId.RABI_FROM_INTENSITY,
    "multipass_2d_rabi": HelperFunctionId.MULTIPASS_2D_RABI,
    "gaussian_beam_rabi": HelperFunctionId.GAUSSIAN_BEAM_RABI,
    "variable_on_off_duty": HelperFunctionId.VARIABLE_ON_OFF_DUTY,
    "variable_on_off_duty_invT": HelperFunctionId.VARIABLE_ON_OFF_DUTY,
    "alternating_sign": HelperFunctionId.ALTERNATING_SIGN,
    "linear_interp": HelperFunctionId.LINEAR_INTERP,
    "pchip_interp": HelperFunctionId.PCHIP_INTERP,
}

HELPER_FUNCTION_ALIASES: Mapping[str, str] = {
    "variable_on_off_duty_invT": "variable_on_off_duty",
}

HELPER_FUNCTION_NAMES: Mapping[int, str] = {
    int(helper_id): name
    for name, helper_id in HELPER_FUNCTION_IDS.items()
    if name not in HELPER_FUNCTION_ALIASES
}
