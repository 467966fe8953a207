"""The runtime-parameter API on the host: the worked examples of the
reference's docstrings (``centrex_tlf/lindblad/parameters.py:229-345``:
``evaluate``, ``evaluate_array``, ``compile_callable`` on a moving-position
profile), checked against closed forms, and the object semantics user code
relies on (immutable parameters compared by identity, one object per name in
an expression, tuple tables, the default parameter set of
``generate_lindblad_parameters``)."""

from __future__ import annotations

import math

import numpy as np
import pytest
import sympy as smp

from centrex_tlf_amd.lindblad.parameters import (
    LindbladParameters,
    Parameter,
    RuntimeExpression,
    Time,
    gaussian,
    generate_lindblad_parameters,
    linear,
    sine,
    tabulated,
)


def test_docstring_examples_against_closed_forms():
    t = Time()
    v = Parameter("v", 200.0)
    z = linear(t, offset=-0.01, slope=v)
    # z = -0.01 + v t: the default velocity, then an override
    assert z.evaluate(t=25e-6) == pytest.approx(-0.01 + 200.0 * 25e-6, rel=1e-15)
    assert z.evaluate(t=25e-6, v=300.0) == pytest.approx(-0.01 + 300.0 * 25e-6, rel=1e-15)
    f = z.compile_callable("t")
    assert f(25e-6) == pytest.approx(-0.005, rel=1e-15)
    # a Gaussian beam crossed at velocity v: Omega(t) = o0 exp(-(z(t) - zl)^2 / (2 w0^2))
    zl, w0, o0 = Parameter("zl", 0.0), Parameter("w0", 1e-3), Parameter("o0", 2.0e6)
    omega = gaussian(z, center=zl, sigma=w0, amplitude=o0)
    ts = np.linspace(0.0, 100e-6, 11)
    want = 2.0e6 * np.exp(-((-0.01 + 200.0 * ts) ** 2) / (2.0 * 1e-3**2))
    np.testing.assert_allclose(omega.evaluate_array("t", ts), want, rtol=1e-13, atol=0)
    np.testing.assert_allclose(omega.evaluate_array("t", ts, o0=1.0), want / 2.0e6, rtol=1e-13, atol=0)


def test_sine_profile_and_unbound_symbols():
    t = Time()
    w = Parameter("w", 3.0)
    s = sine(t, offset=0.5, amplitude=2.0, angular_frequency=w, phase=0.25)
    assert s.evaluate(t=0.1) == pytest.approx(0.5 + 2.0 * math.sin(3.0 * 0.1 + 0.25), rel=1e-15)
    # t left unbound: the expression comes back, with the parameters bound
    left = s.evaluate()
    assert isinstance(left, smp.Expr)
    assert float(left.subs(smp.Symbol("t", real=True), 0.1)) == pytest.approx(s.evaluate(t=0.1), rel=1e-15)


def test_parameters_are_immutable_and_compared_by_identity():
    a, b = Parameter("a", 1.0), Parameter("a", 1.0)
    assert a != b and a == a and hash(a) != hash(b)
    with pytest.raises(AttributeError):
        a.default = 2.0
    with pytest.raises(ValueError, match="duplicate Parameter name 'a'"):
        _ = a + b
    with pytest.raises(ValueError, match="kind must be 'real' or 'complex'"):
        Parameter("k", 1.0, "integer")
    # the phi symbol and the phi letter name one parameter
    assert Parameter("ϕ0").name == Parameter("φ0").name == "φ0"
    assert Parameter("c", 1j, "complex").symbol.is_complex


def test_expression_equality_and_repr():
    v = Parameter("v", 2.0)
    e1, e2 = 3 * v + 1, 3 * v + 1
    assert e1 == e2 and e1 != 3 * v + 2
    with pytest.raises(TypeError):
        hash(e1)
    assert repr(e1).startswith("RuntimeExpression(") and "{v=2.0}" in repr(e1)
    assert repr(RuntimeExpression(smp.Integer(5))) == "RuntimeExpression(5)"
    with pytest.raises(TypeError, match="unsupported runtime expression type str"):
        _ = v + "x"


def test_tuple_table_lookup():
    params = LindbladParameters()
    grid = params.parameter("grid", (0.0, 1.0, 2.0))
    vals = params.parameter("vals", (0.0, 10.0, 40.0))
    x = Parameter("x", 1.5)
    f = tabulated(x, grid, vals)
    assert f.evaluate() == pytest.approx(25.0, rel=1e-15)  # halfway between 10 and 40
    assert f.evaluate(x=0.25) == pytest.approx(2.5, rel=1e-15)
    with pytest.raises(TypeError, match="tabulated requires grid and values"):
        tabulated(x, (0.0, 1.0), vals)


class _Selector:
    def __init__(self, rabi, detuning, polarization, phase_modulation=False):
        self.Ω, self.δ = smp.Symbol(rabi), smp.Symbol(detuning)
        self.polarization_symbols = [smp.Symbol(p) for p in polarization]
        self.phase_modulation = phase_modulation


def test_default_parameter_set_of_transitions():
    gamma = 2.0
    p = generate_lindblad_parameters([_Selector("Ω0", "δ0", ["PZ0"]),
                                      _Selector("Ω1", "δ1", ["PX1", "PY1"], phase_modulation=True)],
                                     gamma=gamma, δ0=0.5)
    base = p.base_parameters
    assert base["Ω0"] == gamma and base["δ0"] == 0.5 and base["PZ0"] == 1.0
    assert base["Ωt1"] == gamma and base["β1"] == 3.8 and base["ωphase1"] == gamma
    assert base["ω1"] == gamma and base["φ1"] == 0.0 and base["δ1"] == 0.0
    comp = p.compound_parameters
    assert set(comp) == {"Ω1", "PX1", "PY1"}
    assert p.parsed_compound_parameters["PY1"] == 1 - smp.Symbol("PX1")
    assert comp["PX1"] == "square_wave(t, ω1, φ1)"
    # compounds are kept after the compounds they read
    order = list(comp)
    assert order.index("PX1") < order.index("PY1")
    with pytest.raises(ValueError, match="more than two symbols"):
        generate_lindblad_parameters([_Selector("Ω0", "δ0", ["A", "B", "C"])])
