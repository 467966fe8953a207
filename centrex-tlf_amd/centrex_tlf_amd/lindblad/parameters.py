"""Runtime parameters of an OBE system: named base values and compound expressions.

Mirrors the user-facing surface of ``centrex_tlf/lindblad/parameters.py``
(``Parameter`` ``:123-176``, ``RuntimeExpression`` ``:179-398``, the waveform
wrappers ``:437-813`` and ``LindbladParameters`` ``:816-1033``):

* a *base* parameter is a number (or a tuple of numbers) that may be scanned
  per trajectory;
* a *compound* parameter is a sympy expression over base parameters, other
  compounds and the time symbol ``t``; it is lowered to the device expression
  bytecode by :mod:`.ir` and re-evaluated on the GPU at every stage when it
  depends on time.

Slots are ordered base parameters first, then compounds in dependency order
(``parameters.py:1000-1033``), which is what the bytecode ``SLOT`` indices
refer to.
"""

from __future__ import annotations

import numbers
from collections import OrderedDict
from collections.abc import Mapping, Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import sympy as smp
from sympy.parsing import sympy_parser

from . import helper_functions as _hf
from .helper_functions import HELPER_FUNCTIONS

__all__ = [
    "GAMMA_DEFAULT",
    "LindbladParameters",
    "Parameter",
    "RuntimeExpression",
    "Time",
    "adapt_lindblad_parameters",
    "alternating_sign",
    "gaussian",
    "gaussian_1d",
    "gaussian_2d",
    "gaussian_2d_rotated",
    "gaussian_beam_rabi",
    "generate_lindblad_parameters",
    "linear",
    "linear_interp",
    "multipass_2d_intensity",
    "multipass_2d_rabi",
    "pchip_interp",
    "pchip_tabulated",
    "phase_modulation",
    "rabi_from_intensity",
    "resonant_polarization_modulation",
    "sawtooth_wave",
    "sine",
    "square_wave",
    "square_wave_profile",
    "tabulated",
    "variable_on_off",
    "variable_on_off_duty",
    "variable_on_off_duty_invT",
]

# Default decay rate of the TlF B state, 1/tau with tau = 99 ns
# (centrex_tlf/constants/constants.py:340-347, 375).
GAMMA_DEFAULT = 1.0 / 99e-9


def _normalize_symbol_name(value: str) -> str:
    # U+03D5 (phi symbol) and U+03C6 (phi letter) name the same parameter.
    return value.replace("ϕ", "φ")


def _is_sequence(value: Any) -> bool:
    if isinstance(value, np.ndarray):
        return value.ndim == 1
    return isinstance(value, Sequence) and not isinstance(value, str | bytes | bytearray)


def _coerce_base_value(value: Any) -> complex | tuple[complex, ...]:
    if isinstance(value, bool):
        return complex(float(value), 0.0)
    if isinstance(value, np.generic):
        return _coerce_base_value(value.item())
    if isinstance(value, numbers.Number):
        return complex(value)
    if _is_sequence(value):
        if isinstance(value, np.ndarray):
            return tuple(_coerce_base_value(v.item()) for v in value.reshape(-1))
        return tuple(_coerce_base_value(v) for v in value)
    raise TypeError(f"unsupported base parameter type {type(value).__name__}")


def _infer_kind(default: Any) -> str:
    coerced = _coerce_base_value(default)
    values = coerced if isinstance(coerced, tuple) else (coerced,)
    return "complex" if any(abs(v.imag) > 0.0 for v in values) else "real"


def _parse_expr(expr: str) -> smp.Expr:
    local_dict = {name: smp.Function(name) for name in HELPER_FUNCTIONS}
    return sympy_parser.parse_expr(_normalize_symbol_name(expr), local_dict=local_dict)


def _merge(left: Mapping[str, "Parameter"], right: Mapping[str, "Parameter"]) -> dict:
    merged = dict(left)
    for name, parameter in right.items():
        existing = merged.get(name)
        if existing is not None and existing is not parameter:
            raise ValueError(f"duplicate Parameter name {name!r}")
        merged[name] = parameter
    return merged


class _Arith:
    """Operator overloads shared by Parameter and RuntimeExpression."""

    def _rt(self) -> "RuntimeExpression":
        raise NotImplementedError

    def __neg__(self):
        rt = self._rt()
        return RuntimeExpression(-rt.expr, rt.parameters)

    def __add__(self, other):
        return _binary(self._rt(), other, lambda a, b: a + b)

    def __radd__(self, other):
        return _binary(_to_runtime_expression(other), self._rt(), lambda a, b: a + b)

    def __sub__(self, other):
        return _binary(self._rt(), other, lambda a, b: a - b)

    def __rsub__(self, other):
        return _binary(_to_runtime_expression(other), self._rt(), lambda a, b: a - b)

    def __mul__(self, other):
        return _binary(self._rt(), other, lambda a, b: a * b)

    def __rmul__(self, other):
        return _binary(_to_runtime_expression(other), self._rt(), lambda a, b: a * b)

    def __truediv__(self, other):
        return _binary(self._rt(), other, lambda a, b: a / b)

    def __rtruediv__(self, other):
        return _binary(_to_runtime_expression(other), self._rt(), lambda a, b: a / b)

    def __pow__(self, other):
        return _binary(self._rt(), other, lambda a, b: a**b)

    def __rpow__(self, other):
        return _binary(_to_runtime_expression(other), self._rt(), lambda a, b: a**b)


@dataclass(frozen=True, eq=False)
class Parameter(_Arith):
    """A registered base parameter that may be scanned or used in expressions."""

    name: str
    default: Any = 0.0
    kind: str = "real"
    symbol: smp.Symbol = field(init=False)

    def __post_init__(self) -> None:
        normalized = _normalize_symbol_name(self.name)
        if self.kind not in {"real", "complex"}:
            raise ValueError("kind must be 'real' or 'complex'")
        object.__setattr__(self, "name", normalized)
        assumptions = {"real": True} if self.kind == "real" else {"complex": True}
        object.__setattr__(self, "symbol", smp.Symbol(normalized, **assumptions))

    def __hash__(self) -> int:
        return id(self)

    def as_expression(self) -> "RuntimeExpression":
        return RuntimeExpression(self.symbol, {self.name: self})

    def _rt(self) -> "RuntimeExpression":
        return self.as_expression()


@dataclass(frozen=True)
class RuntimeExpression(_Arith):
    """A sympy expression plus the registered parameters it depends on."""

    expr: smp.Expr
    parameters: Mapping[str, Parameter] = field(default_factory=dict)

    def __post_init__(self) -> None:
        object.__setattr__(self, "expr", smp.sympify(self.expr))
        object.__setattr__(self, "parameters", dict(self.parameters))

    def _rt(self) -> "RuntimeExpression":
        return self

    def _substituted(self, variable: str | None, overrides: Mapping[str, Any]) -> smp.Expr:
        by_name = {s.name: s for s in self.expr.free_symbols}
        subs = []
        remaining = dict(overrides)
        for name, param in self.parameters.items():
            if name == variable or name not in by_name:
                continue
            if isinstance(param.default, tuple):
                subs.append((by_name[name], smp.Tuple(*param.default)))
            elif name in remaining:
                subs.append((by_name[name], remaining.pop(name)))
            else:
                subs.append((by_name[name], param.default))
        for key, val in remaining.items():
            if key != variable and key in by_name:
                subs.append((by_name[key], val))
        return self.expr.subs(subs)

    def evaluate(self, **overrides: float | complex) -> float | complex:
        """Evaluate with parameter defaults (or ``overrides``), including ``t``."""
        expr = self._substituted(None, overrides)
        if not expr.free_symbols:
            result = smp.lambdify([smp.Symbol("_unused")], expr,
                                  modules=[HELPER_FUNCTIONS, "numpy"])(0)
            if isinstance(result, complex) or np.iscomplexobj(result):
                return complex(result)
            return float(result)
        return complex(expr)

    def evaluate_array(self, variable: str, values: np.ndarray, **overrides) -> np.ndarray:
        expr = self._substituted(variable, overrides)
        sym = {s.name: s for s in self.expr.free_symbols}.get(variable, smp.Symbol(variable))
        func = smp.lambdify(sym, expr, modules=[HELPER_FUNCTIONS, "numpy"])
        return np.asarray(func(np.asarray(values)), dtype=np.float64)

    def compile_callable(self, variable: str = "t", **overrides):
        expr = self._substituted(variable, overrides)
        sym = {s.name: s for s in self.expr.free_symbols}.get(variable, smp.Symbol(variable))
        return smp.lambdify(sym, expr, modules=[HELPER_FUNCTIONS, "numpy"])

    def __repr__(self) -> str:
        return f"RuntimeExpression({self.expr})"


ExpressionLike = RuntimeExpression | Parameter | smp.Expr | int | float | complex


def _to_runtime_expression(value: Any) -> RuntimeExpression:
    if isinstance(value, RuntimeExpression):
        return value
    if isinstance(value, Parameter):
        return value.as_expression()
    if isinstance(value, smp.Expr):
        return RuntimeExpression(value, {})
    if isinstance(value, np.generic):
        return RuntimeExpression(smp.sympify(value.item()), {})
    if isinstance(value, numbers.Number):
        return RuntimeExpression(smp.sympify(value), {})
    raise TypeError(f"unsupported runtime expression type {type(value).__name__}")


def _binary(left: RuntimeExpression, right: Any, op) -> RuntimeExpression:
    rhs = _to_runtime_expression(right)
    return RuntimeExpression(op(left.expr, rhs.expr), _merge(left.parameters, rhs.parameters))


def _call_function(name: str, *args: Any) -> RuntimeExpression:
    coerced = [_to_runtime_expression(arg) for arg in args]
    parameters: dict[str, Parameter] = {}
    for arg in coerced:
        parameters = _merge(parameters, arg.parameters)
    return RuntimeExpression(smp.Function(name)(*(arg.expr for arg in coerced)), parameters)


def _helper_or_value(name: str, *args: Any):
    if any(isinstance(a, Parameter | RuntimeExpression | smp.Expr) for a in args):
        return _call_function(name, *args)
    return HELPER_FUNCTIONS[name](*args)


def Time() -> RuntimeExpression:
    """The integration time ``t`` (``parameters.py:437``)."""
    return RuntimeExpression(smp.Symbol("t", real=True), {})


def gaussian_1d(x, center, sigma):
    return _helper_or_value("gaussian_1d", x, center, sigma)


def gaussian_2d(x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y, theta=None):
    if theta is None:
        return _helper_or_value("gaussian_2d", x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y)
    return gaussian_2d_rotated(x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y, theta)


def gaussian_2d_rotated(x, y, amplitude, mean_x, mean_y, sigma_x, sigma_y, theta):
    return _helper_or_value("gaussian_2d_rotated", x, y, amplitude, mean_x, mean_y,
                            sigma_x, sigma_y, theta)


def phase_modulation(t, beta, omega):
    return _helper_or_value("phase_modulation", t, beta, omega)


def square_wave(t, omega, phase):
    return _helper_or_value("square_wave", t, omega, phase)


def resonant_polarization_modulation(t, gamma, omega):
    return _helper_or_value("resonant_polarization_modulation", t, gamma, omega)


def sawtooth_wave(t, omega, phase):
    return _helper_or_value("sawtooth_wave", t, omega, phase)


def variable_on_off(t, ton, toff, phase):
    return _helper_or_value("variable_on_off", t, ton, toff, phase)


def variable_on_off_duty(t, duty, inv_period, phase):
    return _helper_or_value("variable_on_off_duty", t, duty, inv_period, phase)


def variable_on_off_duty_invT(t, duty, inv_period, phase):
    return _helper_or_value("variable_on_off_duty_invT", t, duty, inv_period, phase)


def multipass_2d_intensity(x, y, amplitudes, xlocs, ylocs, sigma_x, sigma_y):
    return _helper_or_value("multipass_2d_intensity", x, y, amplitudes, xlocs, ylocs,
                            sigma_x, sigma_y)


def rabi_from_intensity(intensity, coupling, dipole_moment=_hf.ED_XTB):
    return _helper_or_value("rabi_from_intensity", intensity, coupling, dipole_moment)


def multipass_2d_rabi(x, y, intensities, xlocs, ylocs, sigma_x, sigma_y, main_coupling,
                      dipole_moment=_hf.ED_XTB):
    return _helper_or_value("multipass_2d_rabi", x, y, intensities, xlocs, ylocs, sigma_x,
                            sigma_y, main_coupling, dipole_moment)


def gaussian_beam_rabi(x, y, intensity, xloc, yloc, sigma_x, sigma_y, main_coupling,
                       dipole_moment=_hf.ED_XTB):
    return _helper_or_value("gaussian_beam_rabi", x, y, intensity, xloc, yloc, sigma_x,
                            sigma_y, main_coupling, dipole_moment)


def alternating_sign(x, x0, width):
    return _helper_or_value("alternating_sign", x, x0, width)


def linear_interp(x, grid, values):
    return _helper_or_value("linear_interp", x, grid, values)


def pchip_interp(x, grid, values):
    return _helper_or_value("pchip_interp", x, grid, values)


def linear(x, offset=0.0, slope=1.0) -> RuntimeExpression:
    return _to_runtime_expression(offset) + (_to_runtime_expression(slope) * _to_runtime_expression(x))


def sine(x, offset=0.0, amplitude=1.0, angular_frequency=1.0, phase=0.0) -> RuntimeExpression:
    arg = _to_runtime_expression(angular_frequency) * _to_runtime_expression(x) + phase
    return _to_runtime_expression(offset) + _to_runtime_expression(amplitude) * RuntimeExpression(
        smp.sin(arg.expr), arg.parameters
    )


def gaussian(x, center=0.0, sigma=1.0, amplitude=1.0, baseline=0.0) -> RuntimeExpression:
    """``baseline + amplitude * exp(-(x - center)^2 / (2 sigma^2))`` (``parameters.py:769-782``)."""
    dx = _to_runtime_expression(x) - center
    width = _to_runtime_expression(sigma)
    exponent = -(dx**2) / (2.0 * width**2)
    return _to_runtime_expression(baseline) + _to_runtime_expression(amplitude) * RuntimeExpression(
        smp.exp(exponent.expr), exponent.parameters
    )


def square_wave_profile(x, low=0.0, high=1.0, period=1.0, phase=0.0, duty=0.5) -> RuntimeExpression:
    on = _call_function("variable_on_off_duty", x, duty, 1.0 / _to_runtime_expression(period), phase)
    return _to_runtime_expression(low) + (_to_runtime_expression(high) - low) * on


def tabulated(x, grid: Parameter, values: Parameter) -> RuntimeExpression:
    if not isinstance(grid, Parameter) or not isinstance(values, Parameter):
        raise TypeError("tabulated requires grid and values to be registered tuple Parameters")
    return _call_function("linear_interp", x, grid, values)


def pchip_tabulated(x, grid: Parameter, values: Parameter) -> RuntimeExpression:
    if not isinstance(grid, Parameter) or not isinstance(values, Parameter):
        raise TypeError("pchip_tabulated requires grid and values to be registered tuple Parameters")
    return _call_function("pchip_interp", x, grid, values)


class LindbladParameters:
    """Base values plus compound expressions bound to Hamiltonian symbols."""

    def __init__(self, base_parameters: Mapping[Any, Any] | None = None,
                 compound_parameters: Mapping[Any, Any] | None = None) -> None:
        self.base_parameters: OrderedDict[str, complex | tuple[complex, ...]] = OrderedDict()
        self.compound_parameters: OrderedDict[str, str] = OrderedDict()
        self._compound_expressions: OrderedDict[str, smp.Expr] = OrderedDict()
        self._parameters_by_name: dict[str, Parameter] = {}
        if base_parameters is not None:
            if compound_parameters is None and self._looks_like_binding_mapping(base_parameters):
                for symbol, expression in base_parameters.items():
                    self.bind(symbol, expression, finalize=False)
            else:
                for name, value in base_parameters.items():
                    self._set_base(str(name), value)
        if compound_parameters is not None:
            for name, expression in compound_parameters.items():
                self._set_compound(str(name), expression)
        self._finalize()

    @staticmethod
    def _looks_like_binding_mapping(mapping: Mapping[Any, Any]) -> bool:
        return any(
            isinstance(key, smp.Symbol) or isinstance(value, RuntimeExpression | Parameter | smp.Expr)
            for key, value in mapping.items()
        )

    @classmethod
    def from_kwargs(cls, **kwargs: Any) -> "LindbladParameters":
        if any(isinstance(v, RuntimeExpression | Parameter | smp.Expr) for v in kwargs.values()):
            return cls(kwargs)
        base, compound = OrderedDict(), OrderedDict()
        for key, value in kwargs.items():
            (compound if isinstance(value, str) else base)[key] = value
        return cls(base, compound)

    def parameter(self, name: str, default: Any = 0.0, kind: str | None = None) -> Parameter:
        parameter = Parameter(name, default, _infer_kind(default) if kind is None else kind)
        self._register_parameter(parameter)
        return parameter

    def real(self, name: str, default: Any = 0.0) -> Parameter:
        return self.parameter(name, default, kind="real")

    def complex(self, name: str, default: Any = 0.0) -> Parameter:
        return self.parameter(name, default, kind="complex")

    def time(self) -> RuntimeExpression:
        return Time()

    def bind(self, symbol: Any, expression: Any, *, overwrite: bool = True,
             finalize: bool = True) -> "LindbladParameters":
        name = _normalize_symbol_name(str(symbol))
        runtime = _to_runtime_expression(expression)
        for parameter in runtime.parameters.values():
            self._register_parameter(parameter)
        if name in self.compound_parameters and not overwrite:
            raise ValueError(f"symbol {name!r} is already bound")
        if str(runtime.expr) == name and name in self.base_parameters:
            self.compound_parameters.pop(name, None)
            self._compound_expressions.pop(name, None)
        else:
            self._set_compound(name, runtime.expr)
        if finalize:
            self._finalize()
        return self

    def drive(self, selector: Any, *, rabi: Any, detuning: Any,
              overwrite: bool = True) -> "LindbladParameters":
        self.bind(selector.Ω, rabi, overwrite=overwrite, finalize=False)
        self.bind(selector.δ, detuning, overwrite=overwrite, finalize=False)
        self._finalize()
        return self

    @property
    def all_parameter_names(self) -> list[str]:
        return list(self.base_parameters) + list(self.compound_parameters)

    @property
    def slot_names(self) -> list[str]:
        return self.all_parameter_names

    @property
    def slot_index_by_name(self) -> dict[str, int]:
        return {name: idx for idx, name in enumerate(self.slot_names)}

    @property
    def parsed_compound_parameters(self) -> OrderedDict[str, smp.Expr]:
        return OrderedDict(self._compound_expressions)

    def validate_symbols(self, symbols: Sequence[str]) -> None:
        allowed = set(self.all_parameter_names)
        missing = sorted({
            _normalize_symbol_name(str(s)) for s in symbols
            if _normalize_symbol_name(str(s)) != "t" and _normalize_symbol_name(str(s)) not in allowed
        })
        if missing:
            raise AssertionError(f"Symbol(s) not defined: {', '.join(missing)}")

    def check_hamiltonian_symbols(self, hamiltonian: smp.Matrix) -> None:
        self.validate_symbols([str(s) for s in hamiltonian.free_symbols])

    def _set_base(self, name: str, value: Any) -> None:
        self.base_parameters[_normalize_symbol_name(name)] = _coerce_base_value(value)

    def _set_compound(self, name: str, expression: Any) -> None:
        normalized = _normalize_symbol_name(name)
        if isinstance(expression, RuntimeExpression):
            expr = expression.expr
        elif isinstance(expression, smp.Expr):
            expr = expression
        elif isinstance(expression, str):
            expr = _parse_expr(expression)
        elif isinstance(expression, numbers.Number):
            expr = smp.sympify(expression)
        else:
            raise TypeError(f"unsupported compound expression type {type(expression).__name__}")
        self.compound_parameters[normalized] = _normalize_symbol_name(str(expr))
        self._compound_expressions[normalized] = expr

    def _register_parameter(self, parameter: Parameter) -> None:
        existing = self._parameters_by_name.get(parameter.name)
        if existing is not None and existing is not parameter:
            raise ValueError(f"duplicate Parameter name {parameter.name!r}")
        self._parameters_by_name[parameter.name] = parameter
        value = _coerce_base_value(parameter.default)
        current = self.base_parameters.get(parameter.name)
        if current is not None and current != value:
            raise ValueError(f"base parameter {parameter.name!r} already has a different default")
        self.base_parameters[parameter.name] = value

    def _finalize(self) -> None:
        defined = set(self.all_parameter_names) | {"t"}
        missing = sorted({
            _normalize_symbol_name(str(s))
            for expr in self._compound_expressions.values()
            for s in expr.free_symbols
            if _normalize_symbol_name(str(s)) not in defined
        })
        if missing:
            raise AssertionError(f"Symbol(s) not defined: {', '.join(missing)}")
        # Order compounds so every compound follows the compounds it reads.
        numeric = set(self.base_parameters) | {"t"}
        unordered = list(self.compound_parameters)
        ordered: list[str] = []
        while unordered:
            progressed = False
            for name in unordered:
                deps = {_normalize_symbol_name(str(s))
                        for s in self._compound_expressions[name].free_symbols}
                if all(d in numeric or d in ordered for d in deps):
                    ordered.append(name)
                    progressed = True
            if not progressed:
                raise ValueError("could not resolve compound parameter dependency order")
            unordered = [name for name in unordered if name not in ordered]
        self.compound_parameters = OrderedDict((k, self.compound_parameters[k]) for k in ordered)
        self._compound_expressions = OrderedDict((k, self._compound_expressions[k]) for k in ordered)


def adapt_lindblad_parameters(parameters: Any) -> LindbladParameters:
    if isinstance(parameters, LindbladParameters):
        return parameters
    if isinstance(parameters, Mapping):
        if any(not isinstance(k, str) or isinstance(v, RuntimeExpression | Parameter | smp.Expr)
               for k, v in parameters.items()):
            return LindbladParameters(parameters)
        return LindbladParameters.from_kwargs(**dict(parameters))
    raise TypeError("parameters must be a LindbladParameters instance or a mapping")


def generate_lindblad_parameters(transition_selectors: Sequence[Any], gamma: float = GAMMA_DEFAULT,
                                 **kwargs: Any) -> LindbladParameters:
    """Default parameter set for transition selectors (``parameters.py:1051-1089``)."""
    entries: list[tuple[str, Any]] = []
    for idx, selector in enumerate(transition_selectors):
        if getattr(selector, "phase_modulation", False):
            entries += [
                (str(selector.Ω), f"Ωt{idx}*phase_modulation(t, β{idx}, ωphase{idx})"),
                (f"Ωt{idx}", gamma), (f"β{idx}", 3.8), (f"ωphase{idx}", gamma),
            ]
        else:
            entries.append((str(selector.Ω), gamma))
        entries.append((str(selector.δ), 0.0))
        symbols = list(getattr(selector, "polarization_symbols", []))
        if len(symbols) == 1:
            entries.append((str(symbols[0]), 1.0))
        elif len(symbols) == 2:
            entries += [
                (f"ω{idx}", gamma), (f"φ{idx}", 0.0),
                (str(symbols[0]), f"square_wave(t, ω{idx}, φ{idx})"),
                (str(symbols[1]), f"1 - {symbols[0]}"),
            ]
        elif len(symbols) > 2:
            raise ValueError("polarization switching with more than two symbols is not supported")
    merged = OrderedDict(entries)
    merged.update(kwargs)
    return adapt_lindblad_parameters(merged)

