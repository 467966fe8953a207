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

import functools
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


@functools.singledispatch
def _slot_value(value: Any) -> complex | tuple[complex, ...]:
    """A base parameter's stored value: a complex scalar, or a tuple of them for
    a one-dimensional sequence (the tuple-valued parameters of the helpers)."""
    if isinstance(value, Sequence) and not isinstance(value, str | bytes | bytearray):
        return tuple(_slot_value(item) for item in value)
    raise TypeError(f"unsupported base parameter type {type(value).__name__}")


@_slot_value.register
def _(value: bool) -> complex:
    return complex(float(value), 0.0)


@_slot_value.register
def _(value: numbers.Number) -> complex:
    return complex(value)


@_slot_value.register
def _(value: np.generic) -> complex | tuple[complex, ...]:
    return _slot_value(value.item())


@_slot_value.register
def _(value: np.ndarray) -> tuple[complex, ...]:
    if value.ndim != 1:
        raise TypeError(f"unsupported base parameter type {type(value).__name__}")
    return tuple(_slot_value(item.item()) for item in value)


def _infer_kind(default: Any) -> str:
    stored = _slot_value(default)
    parts = stored if isinstance(stored, tuple) else (stored,)
    return "complex" if any(part.imag != 0.0 for part in parts) else "real"


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


def _is_binding_mapping(mapping: Mapping[Any, Any]) -> bool:
    """A mapping of Hamiltonian symbols to expressions (rather than names to numbers)."""
    expression_types = RuntimeExpression | Parameter | smp.Expr
    return any(isinstance(k, smp.Symbol) or isinstance(v, expression_types) for k, v in mapping.items())


def _symbol_names(expr: smp.Expr) -> set[str]:
    return {_normalize_symbol_name(str(symbol)) for symbol in expr.free_symbols}


def _greedy_dependency_order(names: Sequence[str], needs: Mapping[str, set[str]], known: set[str]) -> list[str]:
    """Names in sweeps over the pending list, each name placed as soon as every
    name it needs is known or placed (the slot order of parameters.py:1000-1033)."""
    placed: list[str] = []
    resolved = set(known)
    pending = list(names)
    while pending:
        before = len(placed)
        for name in pending:
            if needs[name] <= resolved:
                placed.append(name)
                resolved.add(name)
        if len(placed) == before:
            raise ValueError("could not resolve compound parameter dependency order")
        pending = [name for name in pending if name not in resolved]
    return placed


class LindbladParameters:
    """Base values plus compound expressions bound to Hamiltonian symbols.

    ``base_parameters`` maps names to stored values (complex, or tuples for
    table-valued parameters); ``compound_parameters`` maps bound symbols to
    their expression text, kept in slot order after every change.
    """

    def __init__(self, base_parameters: Mapping[Any, Any] | None = None,
                 compound_parameters: Mapping[Any, Any] | None = None) -> None:
        self.base_parameters: OrderedDict[str, complex | tuple[complex, ...]] = OrderedDict()
        self.compound_parameters: OrderedDict[str, str] = OrderedDict()
        self._exprs: OrderedDict[str, smp.Expr] = OrderedDict()
        self._registry: dict[str, Parameter] = {}
        if base_parameters is not None and compound_parameters is None and _is_binding_mapping(base_parameters):
            for symbol, expression in base_parameters.items():
                self.bind(symbol, expression, finalize=False)
        else:
            for name, value in (base_parameters or {}).items():
                self._store_base(str(name), value)
            for name, expression in (compound_parameters or {}).items():
                self._store_compound(str(name), expression)
        self._finalize()

    @classmethod
    def from_kwargs(cls, **kwargs: Any) -> "LindbladParameters":
        """Keyword form: strings are compound expressions, numbers base values;
        any expression object makes the whole set a binding mapping."""
        if _is_binding_mapping(kwargs):
            return cls(kwargs)
        texts = OrderedDict((k, v) for k, v in kwargs.items() if isinstance(v, str))
        values = OrderedDict((k, v) for k, v in kwargs.items() if not isinstance(v, str))
        return cls(values, texts)

    def parameter(self, name: str, default: Any = 0.0, kind: str | None = None) -> Parameter:
        created = Parameter(name, default, _infer_kind(default) if kind is None else kind)
        self._register_parameter(created)
        return created

    def real(self, name: str, default: Any = 0.0) -> Parameter:
        return self.parameter(name, default, kind="real")

    def complex(self, name: str, default: Any = 0.0) -> Parameter:
        return self.parameter(name, default, kind="complex")

    def time(self) -> RuntimeExpression:
        return Time()

    def bind(self, symbol: Any, expression: Any, *, overwrite: bool = True,
             finalize: bool = True) -> "LindbladParameters":
        """Bind a Hamiltonian symbol to an expression (registering the
        expression's parameters); binding a base parameter to itself removes
        an earlier compound of that name."""
        target = _normalize_symbol_name(str(symbol))
        rhs = _to_runtime_expression(expression)
        for used in rhs.parameters.values():
            self._register_parameter(used)
        if not overwrite and target in self.compound_parameters:
            raise ValueError(f"symbol {target!r} is already bound")
        is_self_binding = target in self.base_parameters and str(rhs.expr) == target
        if is_self_binding:
            self.compound_parameters.pop(target, None)
            self._exprs.pop(target, None)
        else:
            self._store_compound(target, rhs.expr)
        if finalize:
            self._finalize()
        return self

    def drive(self, selector: Any, *, rabi: Any, detuning: Any,
              overwrite: bool = True) -> "LindbladParameters":
        """Bind a transition's Rabi rate and detuning symbols at once."""
        for symbol, expression in ((selector.Ω, rabi), (selector.δ, detuning)):
            self.bind(symbol, expression, overwrite=overwrite, finalize=False)
        self._finalize()
        return self

    @property
    def all_parameter_names(self) -> list[str]:
        return [*self.base_parameters, *self.compound_parameters]

    @property
    def slot_names(self) -> list[str]:
        return self.all_parameter_names

    @property
    def slot_index_by_name(self) -> dict[str, int]:
        return dict(zip(self.slot_names, range(len(self.slot_names))))

    @property
    def parsed_compound_parameters(self) -> OrderedDict[str, smp.Expr]:
        return OrderedDict(self._exprs)

    def validate_symbols(self, symbols: Sequence[str]) -> None:
        known = set(self.all_parameter_names) | {"t"}
        unknown = {_normalize_symbol_name(str(s)) for s in symbols} - known
        if unknown:
            raise AssertionError(f"Symbol(s) not defined: {', '.join(sorted(unknown))}")

    def check_hamiltonian_symbols(self, hamiltonian: smp.Matrix) -> None:
        self.validate_symbols([str(s) for s in hamiltonian.free_symbols])

    # ---- storage

    def _store_base(self, name: str, value: Any) -> None:
        self.base_parameters[_normalize_symbol_name(name)] = _slot_value(value)

    def _store_compound(self, name: str, expression: Any) -> None:
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
        key = _normalize_symbol_name(name)
        self._exprs[key] = expr
        self.compound_parameters[key] = _normalize_symbol_name(str(expr))

    def _register_parameter(self, parameter: Parameter) -> None:
        """Record a Parameter object as a base slot; one object per name, and
        one default per name."""
        if self._registry.setdefault(parameter.name, parameter) is not parameter:
            raise ValueError(f"duplicate Parameter name {parameter.name!r}")
        value = _slot_value(parameter.default)
        if self.base_parameters.setdefault(parameter.name, value) != value:
            raise ValueError(f"base parameter {parameter.name!r} already has a different default")

    def _finalize(self) -> None:
        """Check every compound's symbols are defined, then put the compounds in
        slot order (each after the compounds it reads)."""
        known = set(self.base_parameters) | {"t"}
        needs = {name: _symbol_names(expr) for name, expr in self._exprs.items()}
        undefined = set().union(*needs.values()) - known - set(self._exprs)
        if undefined:
            raise AssertionError(f"Symbol(s) not defined: {', '.join(sorted(undefined))}")
        order = _greedy_dependency_order(list(self.compound_parameters), needs, known)
        self.compound_parameters = OrderedDict((name, self.compound_parameters[name]) for name in order)
        self._exprs = OrderedDict((name, self._exprs[name]) for name in order)


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

