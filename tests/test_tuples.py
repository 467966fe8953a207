"""Tuple-valued parameters: the oracle's tuple helpers pinned by closed forms,
and the host lowering of tuples to scalar programs checked against the
tuple-valued evaluation it replaces.

Helpers (rust/src/lindblad/eval.rs:250-451, plan.rs:33-115), evaluated by the
oracle's machine with their tuples passed element by element:

- linear_interp (14) is exact on linear data and flat outside the grid;
- pchip_interp (16, the Hermite form) equals scipy's PchipInterpolator on
  intervals that touch no end knot (scipy's end slopes are a different
  three-point rule; the interior slopes are the same weighted harmonic mean),
  reproduces the knots, reproduces linear data everywhere and is monotone on
  monotone data;
- the table form (17) is the same cubic in Horner form: it agrees with 16 to
  rounding at every point;
- multipass_2d_intensity (8) is a sum of separable Gaussians (scipy's normal
  densities), multipass_2d_rabi (10) is E = sqrt(2 I / (c eps0)) times
  coupling times dipole / hbar with scipy.constants;
- the reference's own numeric cases (tests/lindblad/
  test_parameter_helper_wrappers.py:77-106).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.constants as sc
import sympy as smp
from scipy.interpolate import PchipInterpolator
from scipy.stats import norm

from centrex_tlf_amd.lindblad import LindbladParameters, Time, prepare_lindblad_problem, two_level_system
from centrex_tlf_amd.lindblad import parameters as P
from centrex_tlf_amd.lindblad.helper_functions import ED_XTB
from centrex_tlf_amd.lindblad.ir import evaluate_parameter_graph
from centrex_tlf_amd.lindblad.parameters import RuntimeExpression
from centrex_tlf_amd.lindblad.tuples import extract_pchip_tables
from oracle.lb_oracle import Oracle, eval_expression

CONST, TIME, HELPER = 1, 4, 13


def tuple_helper(fid: int, x: float, *groups) -> float:
    """Helper `fid` at x with scalars and tuples in `groups` passed element by element."""
    args = []
    for g in groups:
        args += list(g) if isinstance(g, (tuple, list, np.ndarray)) else [g]
    prog = ([(TIME, 0, 0, 0, 0.0)] + [(CONST, 0, 0, 0, complex(a)) for a in args]
            + [(HELPER, 0, 1 + len(args), fid, 0.0)])
    v = eval_expression(prog, x)
    assert v.imag == 0.0
    return v.real


GRID = np.array([0.0, 0.4, 1.1, 1.5, 2.6, 3.0, 4.2])
VALS = np.array([0.3, 1.2, 0.8, 0.85, 2.0, -0.4, 0.1])


def test_linear_interp_is_exact_on_linear_data():
    values = 1.7 * GRID - 0.6
    for x in np.linspace(-0.5, 4.7, 211):
        want = 1.7 * min(max(x, GRID[0]), GRID[-1]) - 0.6
        assert abs(tuple_helper(14, x, GRID, values) - want) <= 1e-14 * (1 + abs(want)), x
    assert tuple_helper(14, 9.0, (2.0,), (5.5,)) == 5.5  # one point: its value everywhere


def test_pchip_is_scipy_on_interior_intervals_and_hits_the_knots():
    ref = PchipInterpolator(GRID, VALS)
    xs = np.linspace(GRID[1], GRID[-2], 301)
    got = np.array([tuple_helper(16, x, GRID, VALS) for x in xs])
    np.testing.assert_allclose(got, ref(xs), rtol=0, atol=1e-13)
    for g, v in zip(GRID, VALS):
        assert abs(tuple_helper(16, g, GRID, VALS) - v) <= 1e-15 * (1 + abs(v))
    assert tuple_helper(16, -3.0, GRID, VALS) == VALS[0]
    assert tuple_helper(16, 9.0, GRID, VALS) == VALS[-1]


def test_pchip_reproduces_lines_and_keeps_monotone_data_monotone():
    lin = -0.8 * GRID + 2.0
    for x in np.linspace(GRID[0], GRID[-1], 97):
        assert abs(tuple_helper(16, x, GRID, lin) - (-0.8 * x + 2.0)) <= 1e-14
    mono = np.cumsum([0.1, 2.0, 0.0, 0.05, 1.5, 0.3, 0.01])
    ys = [tuple_helper(16, x, GRID, mono) for x in np.linspace(GRID[0], GRID[-1], 401)]
    assert np.all(np.diff(ys) >= -1e-15)


def test_pchip_table_form_is_the_hermite_form_to_rounding():
    for x in np.concatenate([np.linspace(-0.3, 4.5, 241), GRID]):
        a, b = tuple_helper(16, x, GRID, VALS), tuple_helper(17, x, GRID, VALS)
        assert abs(a - b) <= 4e-15 * (1 + abs(a)), x


def test_multipass_helpers_are_sums_of_gaussian_spots():
    amps, xl, yl = (1.0, 0.5, 2.2), (-0.3, 0.4, 0.05), (0.25, -0.1, 0.6)
    x, y, sx, sy = 0.1, -0.2, 0.7, 0.9
    # amplitude * exp(-dx^2/2sx^2 - dy^2/2sy^2) = amplitude * 2 pi sx sy * N(x; xl, sx) N(y; yl, sy)
    want = sum(a * 2 * math.pi * sx * sy * norm.pdf(x, xc, sx) * norm.pdf(y, yc, sy)
               for a, xc, yc in zip(amps, xl, yl))
    prog_args = [y, amps, xl, yl, sx, sy]
    got = tuple_helper(8, x, *prog_args)
    assert abs(got - want) <= 1e-14 * want
    coupling = 0.35
    field = math.sqrt(2.0 * want / (sc.c * sc.epsilon_0))
    dipole = 0.315 * sc.e * sc.physical_constants["Bohr radius"][0]
    rabi = field * coupling * dipole / sc.hbar
    assert abs(tuple_helper(10, x, *prog_args, coupling) - rabi) <= 1e-9 * rabi
    assert abs(tuple_helper(10, x, *prog_args, coupling, 2.0 * dipole) - 2.0 * rabi) <= 1e-9 * rabi
    assert abs(ED_XTB - dipole) <= 1e-9 * dipole


def test_the_reference_numeric_cases():
    # tests/lindblad/test_parameter_helper_wrappers.py:77-106
    assert tuple_helper(14, 0.25, (0.0, 1.0), (2.0, 4.0)) == pytest.approx(2.5, abs=1e-15)
    assert tuple_helper(16, 0.25, (0.0, 0.5, 1.0), (2.0, 3.0, 4.0)) == pytest.approx(2.5, abs=1e-15)
    dx = np.array([0.1 + 0.3, 0.1 - 0.4])
    dy = np.array([-0.2 - 0.25, -0.2 + 0.1])
    want = float(np.sum(np.array([1.0, 0.5]) * np.exp(-(dx ** 2 / (2 * 0.49) + dy ** 2 / (2 * 0.81)))))
    got = tuple_helper(8, 0.1, -0.2, (1.0, 0.5), (-0.3, 0.4), (0.25, -0.1), 0.7, 0.9)
    assert got == pytest.approx(want, rel=1e-15)


# ---- the host lowering -----------------------------------------------------

def _tuple_problem():
    p = LindbladParameters()
    t = Time()
    grid = p.real("grid", tuple(GRID))
    vals = p.real("vals", tuple(VALS))
    amps = p.real("amps", (1.0, 0.5, 0.25))
    xl = p.real("xl", (-0.3, 0.4, 0.1))
    yl = p.real("yl", (0.25, -0.1, 0.0))
    w = p.real("w", 1.3)
    p.bind("scaled", amps * w, finalize=False)  # a tuple-valued compound
    scaled = RuntimeExpression(smp.Symbol("scaled"), {})
    p.bind("Omega", P.pchip_interp(t, grid, vals) * w
           + P.multipass_2d_intensity(0.1, -0.2, scaled, xl, yl, 0.7, 0.9), finalize=False)
    p.bind("delta", P.linear_interp(t * 0.5, grid, vals) - w, finalize=False)
    p._finalize()
    system = two_level_system(omega_name="Omega", delta_name="delta")
    return prepare_lindblad_problem(system, p, backend="host")


def test_lowered_programs_evaluate_as_the_tuple_graph():
    prepared = _tuple_problem()
    flat = prepared.flat_plan
    assert 17 in set(flat["instr_function"][flat["instr_op"] == HELPER])  # the table form was matched
    oracle = Oracle(prepared)
    # the same system with Omega and delta as plain parameters
    q = LindbladParameters()
    q.bind("Omega", q.complex("Om", 0.0), finalize=False)
    q.bind("delta", q.complex("de", 0.0), finalize=False)
    q._finalize()
    plain = Oracle(prepare_lindblad_problem(two_level_system(omega_name="Omega", delta_name="delta"), q,
                                            backend="host"))
    slot_of = {n: k for k, n in enumerate(prepared.slot_names)}
    rho = np.array([[0.6, 0.2 + 0.1j], [0.2 - 0.1j, 0.4]])
    y = prepared.layout.pack(rho)
    amps, xl, yl = np.array([1.0, 0.5, 0.25]), (-0.3, 0.4, 0.1), (0.25, -0.1, 0.0)
    for t in np.linspace(-0.2, 4.5, 23):
        for w in (1.3, 0.7):
            om = (tuple_helper(17, t, GRID, VALS) * w
                  + tuple_helper(8, 0.1, -0.2, tuple(amps * w), xl, yl, 0.7, 0.9))
            de = tuple_helper(14, 0.5 * t, GRID, VALS) - w
            if GRID[1] <= t <= GRID[-2]:  # the host graph (scipy's pchip) agrees inside the end intervals
                slots = evaluate_parameter_graph(prepared.parameter_graph, t, {slot_of["w"]: w})
                assert abs(complex(slots[slot_of["Omega"]]) - om) <= 1e-12 * (1 + abs(om))
                assert abs(complex(slots[slot_of["delta"]]) - de) <= 1e-14 * (1 + abs(de))
            want = plain.rhs(y, t, {"Om": om, "de": de})
            got = oracle.rhs(y, t, {"w": w})
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-13, err_msg=f"t={t} w={w}")


def test_pchip_tables_pair_grids_with_every_same_length_tuple():
    """The reference's table extraction and grid-only match
    (centrex_tlf/lindblad/ir.py:361-397, rust/src/lindblad/eval.rs:590-626):
    with two value tuples of the grid's length, a call on the second one
    evaluates the first one's table."""
    base = [(0.0, 1.0, 2.0), (1.0, 3.0, 2.0), 0.5, (5.0, 4.0, 3.0)]
    tables = extract_pchip_tables(base)
    assert tables[0] == ((0.0, 1.0, 2.0), (1.0, 3.0, 2.0))
    assert tables[1] == ((0.0, 1.0, 2.0), (5.0, 4.0, 3.0))
    p = LindbladParameters()
    g = p.real("g", base[0])
    v1 = p.real("v1", base[1])
    p.real("s", 0.5)
    v2 = p.real("v2", base[3])
    p.bind("Omega", P.pchip_interp(Time(), g, v2), finalize=False)
    p.bind("delta", 0.0, finalize=False)
    p._finalize()
    prepared = prepare_lindblad_problem(two_level_system(omega_name="Omega", delta_name="delta"), p,
                                        backend="host")
    flat = prepared.flat_plan
    h = np.flatnonzero((flat["instr_op"] == HELPER) & (flat["instr_function"] == 17))
    assert len(h) == 1
    consts = flat["instr_re"][h[0] - 6:h[0]]
    np.testing.assert_array_equal(consts, list(base[0]) + list(base[1]))  # v1's table, not v2
    del v1


@pytest.mark.parametrize("build,message", [
    (lambda p, t: P.linear_interp(t, p.real("g", (0.0, 1.0)), p.real("v", (1.0, 2.0, 3.0))),
     "same length"),
    (lambda p, t: P.linear_interp(t, p.real("g", (0.0, 0.0)), p.real("v", (1.0, 2.0))), "strictly increasing"),
    (lambda p, t: P.sine(p.real("g", (0.0, 1.0))), "sequence value"),
])
def test_lowering_rejects_what_the_reference_rejects(build, message):
    p = LindbladParameters()
    p.bind("Omega", build(p, Time()), finalize=False)
    p.bind("delta", 0.0, finalize=False)
    p._finalize()
    with pytest.raises((ValueError, TypeError), match=message):
        prepare_lindblad_problem(two_level_system(omega_name="Omega", delta_name="delta"), p, backend="host")
