"""Effective-Hamiltonian test plans built from dense superoperators.

Two levels |0>, |1>: H_int(x) = -kappa * x * |1><1| (or -kappa x^2 for the
pchip checks), decay |1> -> |0> at rate 0.3, H_opt = sigma_x (scaled by
Omega/2 in the RHS), H_det = -|1><1| (scaled by delta).  The field
coordinate is a compound slot x(t) = x0 + v t, so the dense Lindblad
superoperator at x(t) is known in closed form and the same physics runs on
the main Lindblad path as H = [[0, Omega/2], [Omega/2, -delta - kappa x(t)]].
Also a random n-level plan with sparse random operators for the GPU checks.
"""

from __future__ import annotations

import numpy as np
import sympy as smp

from centrex_tlf_amd.lindblad import LindbladParameters, Time, gaussian
from centrex_tlf_amd.lindblad.effective import effective_operator_payload, split_real_superoperator
from centrex_tlf_amd.lindblad.systems import OBESystem

KAPPA = 0.7
DECAY = 0.3


def commutator_superop(H: np.ndarray) -> np.ndarray:
    """-i[H, rho] on rho.flat (row-major): -i (H (x) I - I (x) H^T)."""
    n = H.shape[0]
    eye = np.eye(n)
    return -1j * (np.kron(H, eye) - np.kron(eye, H.T))


def dissipator_superop(Cop: np.ndarray) -> np.ndarray:
    """C rho C^dag - {C^dag C, rho}/2 on rho.flat."""
    n = Cop.shape[0]
    eye = np.eye(n)
    CdC = Cop.conj().T @ Cop
    return np.kron(Cop, Cop.conj()) - 0.5 * (np.kron(CdC, eye) + np.kron(eye, CdC.T))


def two_level_parameters(x0: float = -1.0, v: float = 0.5, omega: float = 1.1, delta: float = 0.2):
    params = LindbladParameters()
    params.real("Omega", omega)
    params.real("delta", delta)
    x0p = params.real("x0", x0)
    vp = params.real("v", v)
    params.bind("field", x0p + vp * Time(), finalize=False)
    params._finalize()
    return params


def two_level_operators(grid: np.ndarray, quadratic: bool = False):
    Cop = np.zeros((2, 2), dtype=complex)
    Cop[0, 1] = np.sqrt(DECAY)
    comb = []
    for x in grid:
        shift = KAPPA * (x * x if quadratic else x)
        H = np.diag([0.0, -shift]).astype(complex)
        comb.append(split_real_superoperator(commutator_superop(H) + dissipator_superop(Cop)))
    opt = split_real_superoperator(commutator_superop(np.array([[0, 1], [1, 0]], dtype=complex)))
    det = split_real_superoperator(commutator_superop(np.diag([0.0, -1.0]).astype(complex)))
    return comb, [opt] * len(grid), [det] * len(grid)


def two_level_plan_dict(grid=None, interpolation: str = "linear", quadratic: bool = False, **pkw):
    grid = np.linspace(-2.0, 2.0, 9) if grid is None else np.asarray(grid, dtype=float)
    params = two_level_parameters(**pkw)
    comb, opt, det = two_level_operators(grid, quadratic)
    names = list(params.slot_names)
    return {
        "n_states": 2, "real_dim": 8, "n_grid": grid.size, "field_grid": grid,
        "sparse_combined": effective_operator_payload(comb, grid, interpolation),
        "sparse_opt": effective_operator_payload(opt, grid, interpolation),
        "sparse_det": effective_operator_payload(det, grid, interpolation),
        "excited_indices": np.array([1]), "ground_indices": np.array([0]), "sink_indices": np.array([], dtype=int),
        "field_coordinate_slot": names.index("field"), "rabi_rate_slot": names.index("Omega"),
        "detuning_slot": names.index("delta"), "is_time_dependent": True,
        "operator_interpolation": interpolation, "parameter_graph": params,
    }


def dense_rhs(y: np.ndarray, t: float, omega: float = 1.1, delta: float = 0.2, x0: float = -1.0, v: float = 0.5,
              quadratic: bool = False) -> np.ndarray:
    """The split-real Lindblad RHS at t, straight from the matrices."""
    x = x0 + v * t
    shift = KAPPA * (x * x if quadratic else x)
    H = np.array([[0.0, omega / 2], [omega / 2, -delta - shift]], dtype=complex)
    Cop = np.zeros((2, 2), dtype=complex)
    Cop[0, 1] = np.sqrt(DECAY)
    return split_real_superoperator(commutator_superop(H) + dissipator_superop(Cop)) @ y


def two_level_obe(x0: float = -1.0, v: float = 0.5):
    """The same physics for the main Lindblad path."""
    om, de, xx, vv, t = smp.symbols("Omega delta x0 v t", real=True)
    field = xx + vv * t
    H = smp.Matrix([[0, om / 2], [om / 2, -de - KAPPA * field]])
    Cop = np.zeros((1, 2, 2), dtype=complex)
    Cop[0, 0, 1] = np.sqrt(DECAY)
    system = OBESystem(H_symbolic=H, C_array=Cop, QN=[{"index": 0}, {"index": 1}], description="field-shifted")
    return system


def random_plan_dict(n: int = 4, n_grid: int = 6, seed: int = 0, interpolation: str = "pchip",
                     density: float = 0.15):
    """Random sparse operators on a field grid, x(t) = x0 + v t, Omega(t) a
    Gaussian pulse; entries scaled to a spectral radius of about one and the
    combined operator shifted by -1.5 so the state stays bounded."""
    rng = np.random.default_rng(seed)
    D = 2 * n * n
    grid = np.sort(rng.uniform(-1.0, 1.0, n_grid))
    grid[0], grid[-1] = -1.0, 1.0
    norm = 1.0 / np.sqrt(density * D)

    def mats(scale):
        mask = rng.random((D, D)) < density
        base = rng.normal(size=(D, D)) * mask * norm
        slope = rng.normal(size=(D, D)) * mask * norm
        curv = rng.normal(size=(D, D)) * mask * norm
        return [scale * (base + slope * g + 0.3 * curv * g * g) - 1.5 * np.eye(D) * (scale > 0.5) for g in grid]

    params = LindbladParameters()
    params.real("Omega0", 0.8)
    params.real("delta", 0.3)
    x0p = params.real("x0", -0.9)
    vp = params.real("v", 0.35)
    params.bind("field", x0p + vp * Time(), finalize=False)
    om = params.real("Omega_amp", 1.0)
    params.bind("Omega", gaussian(Time(), center=2.5, sigma=1.0, amplitude=om), finalize=False)
    params._finalize()
    names = list(params.slot_names)
    return {
        "n_states": n, "real_dim": D, "n_grid": n_grid, "field_grid": grid,
        "sparse_combined": effective_operator_payload(mats(1.0), grid, interpolation),
        "sparse_opt": effective_operator_payload(mats(0.3), grid, interpolation),
        "sparse_det": effective_operator_payload(mats(0.2), grid, interpolation),
        "excited_indices": np.array([n - 1]), "ground_indices": np.arange(n - 1), "sink_indices": np.array([], dtype=int),
        "field_coordinate_slot": names.index("field"), "rabi_rate_slot": names.index("Omega"),
        "detuning_slot": names.index("delta"), "is_time_dependent": True,
        "operator_interpolation": interpolation, "parameter_graph": params,
    }
