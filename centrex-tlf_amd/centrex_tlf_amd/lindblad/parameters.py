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
import operator
from collections import OrderedDict
from collections.abc import Mapping, Sequence
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


# ---- stored base values


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


# helper names parse as undefined functions, so the lowering sees the call
_HELPER_CALLS = {name: smp.Function(name) for name in HELPER_FUNCTIONS}


def _parse_expr(text: str) -> smp.Expr:
    return sympy_parser.parse_expr(_normalize_symbol_name(text), local_dict=dict(_HELPER_CALLS))


# ---- expression objects


def _freeze(obj: Any, **attributes: Any) -> None:
    for key, value in attributes.items():
        object.__setattr__(obj, key, value)


def _joined(*maps: Mapping[str, "Parameter"]) -> dict[str, "Parameter"]:
    """The union of the parameter maps of several operands; a name must stand
    for one Parameter object throughout an expression."""
    union: dict[str, Parameter] = {}
    for mapping in maps:
        for key, param in mapping.items():
            if union.setdefault(key, param) is not param:
                raise ValueError(f"duplicate Parameter name {key!r}")
    return union


def _combine(build, *operands: Any) -> "RuntimeExpression":
    """``build`` applied to the sympy forms of ``operands`` (each coerced to a
    RuntimeExpression), carrying the union of their parameters."""
    terms = [_to_runtime_expression(item) for item in operands]
    return RuntimeExpression(build(*(term.expr for term in terms)), _joined(*(term.parameters for term in terms)))


class _Immutable:
    __slots__ = ()

    def __setattr__(self, key: str, value: Any) -> None:
        raise AttributeError(f"cannot assign to field {key!r} of an immutable {type(self).__name__}")

    def __delattr__(self, key: str) -> None:
        raise AttributeError(f"cannot delete field {key!r} of an immutable {type(self).__name__}")


def _with_arithmetic(cls):
    """Install +, -, *, /, ** (both operand orders) and unary minus; every
    result is a RuntimeExpression over the joined parameters."""
    for dunder, op in (("add", operator.add), ("sub", operator.sub), ("mul", operator.mul),
                       ("truediv", operator.truediv), ("pow", operator.pow)):
        setattr(cls, f"__{dunder}__", lambda self, other, op=op: _combine(op, self, other))
        setattr(cls, f"__r{dunder}__", lambda self, other, op=op: _combine(op, other, self))
    cls.__neg__ = lambda self: _combine(operator.neg, self)
    return cls


@_with_arithmetic
class Parameter(_Immutable):
    """A registered base parameter that may be scanned or used in expressions.

    ``Parameter(name, default=0.0, kind="real")``; ``symbol`` is the sympy
    symbol of the normalised name with the kind's assumption.  Two Parameter
    objects are equal only when they are the same object.
    """

    __slots__ = ("name", "default", "kind", "symbol")

    def __init__(self, name: str, default: Any = 0.0, kind: str = "real") -> None:
        key = _normalize_symbol_name(name)
        if kind not in ("real", "complex"):
            raise ValueError("kind must be 'real' or 'complex'")
        _freeze(self, name=key, default=default, kind=kind, symbol=smp.Symbol(key, **{kind: True}))

    __hash__ = object.__hash__

    def __repr__(self) -> str:
        return f"Parameter(name={self.name!r}, default={self.default!r}, kind={self.kind!r})"

    def as_expression(self) -> "RuntimeExpression":
        return RuntimeExpression(self.symbol, {self.name: self})


@_with_arithmetic
class RuntimeExpression(_Immutable):
    """A sympy expression plus the registered parameters it depends on."""

    __slots__ = ("expr", "parameters")

    def __init__(self, expr: Any, parameters: Mapping[str, Parameter] | None = None) -> None:
        _freeze(self, expr=smp.sympify(expr), parameters=dict(parameters or {}))

    def __eq__(self, other: Any) -> bool:
        if type(other) is not RuntimeExpression:
            return NotImplemented
        return self.expr == other.expr and self.parameters == other.parameters

    __hash__ = None  # the parameter map is mutable

    def _bound(self, free: str | None, overrides: Mapping[str, Any]) -> smp.Expr:
        """The expression with every symbol but ``free`` replaced: a tuple
        parameter by its table (overrides do not apply to tables), a scalar
        parameter by its override or default, any other symbol (``t``) by an
        override of its name."""
        values: dict[str, Any] = dict(overrides)
        for key, param in self.parameters.items():
            if isinstance(param.default, tuple):
                values[key] = smp.Tuple(*param.default)
            else:
                values.setdefault(key, param.default)
        values.pop(free, None)
        pairs = [(sym, values[sym.name]) for sym in self.expr.free_symbols if sym.name in values]
        return self.expr.subs(pairs)

    def _variable(self, name: str) -> smp.Symbol:
        return next((sym for sym in self.expr.free_symbols if sym.name == name), smp.Symbol(name))

    def evaluate(self, **overrides: float | complex) -> float | complex:
        """Evaluate with parameter defaults (or ``overrides``), including ``t``.
        An expression with symbols left unbound is returned as it stands."""
        bound = self._bound(None, overrides)
        if bound.free_symbols:
            try:
                return complex(bound)
            except (TypeError, ValueError):
                return bound
        value = smp.lambdify([], bound, modules=[HELPER_FUNCTIONS, "numpy"])()
        return complex(value) if np.iscomplexobj(value) else float(value)

    def compile_callable(self, variable: str = "t", **overrides: float | complex):
        """A numpy function of ``variable`` with every other symbol bound."""
        return smp.lambdify(self._variable(variable), self._bound(variable, overrides),
                            modules=[HELPER_FUNCTIONS, "numpy"])

    def evaluate_array(self, variable: str, values: np.ndarray, **overrides) -> np.ndarray:
        return np.asarray(self.compile_callable(variable, **overrides)(np.asarray(values)), dtype=np.float64)

    def __repr__(self) -> str:
        scalars = ", ".join(f"{key}={param.default}" for key, param in self.parameters.items()
                            if not isinstance(param.default, tuple))
        return f"RuntimeExpression({self.expr}, {{{scalars}}})" if scalars else f"RuntimeExpression({self.expr})"


ExpressionLike = RuntimeExpression | Parameter | smp.Expr | int | float | complex


def _to_runtime_expression(value: Any) -> RuntimeExpression:
    match value:
        case RuntimeExpression():
            return value
        case Parameter():
            return value.as_expression()
        case smp.Expr():
            return RuntimeExpression(value)
        case np.generic():
            return RuntimeExpression(smp.sympify(value.item()))
        case numbers.Number():
            return RuntimeExpression(smp.sympify(value))
    raise TypeError(f"unsupported runtime expression type {type(value).__name__}")


def _call_function(name: str, *args: Any) -> RuntimeExpression:
    """``name(args...)`` as an undefined sympy function (a helper call)."""
    return _combine(smp.Function(name), *args)


def _helper_or_value(name: str, *args: Any):
    """A helper call: symbolic when any argument is an expression, else the
    helper's value."""
    symbolic = (Parameter, RuntimeExpression, smp.Expr)
    return _call_function(name, *args) if any(isinstance(a, symbolic) for a in args) else HELPER_FUNCTIONS[name](*args)


def Time() -> RuntimeExpression:
    """The integration time ``t`` (``parameters.py:437``)."""
    return RuntimeExpression(smp.Symbol("t", real=True))


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


# Profiles built from sympy operations.  The operations run in the order the
# formulas are written, so the stored expression trees (and the bytecode
# lowered from them) do not depend on how a profile was assembled.


def linear(x, offset=0.0, slope=1.0) -> RuntimeExpression:
    """``offset + slope * x``."""
    return _combine(lambda off, k, xv: off + k * xv, offset, slope, x)


def sine(x, offset=0.0, amplitude=1.0, angular_frequency=1.0, phase=0.0) -> RuntimeExpression:
    """``offset + amplitude * sin(angular_frequency * x + phase)``."""
    return _combine(lambda off, amp, w, xv, ph: off + amp * smp.sin(w * xv + ph),
                    offset, amplitude, angular_frequency, x, phase)


def gaussian(x, center=0.0, sigma=1.0, amplitude=1.0, baseline=0.0) -> RuntimeExpression:
    """``baseline + amplitude * exp(-(x - center)^2 / (2 sigma^2))`` (``parameters.py:769-782``)."""
    return _combine(lambda base, amp, xv, mu, sd: base + amp * smp.exp(-((xv - mu) ** 2) / (2.0 * sd**2)),
                    baseline, amplitude, x, center, sigma)


def square_wave_profile(x, low=0.0, high=1.0, period=1.0, phase=0.0, duty=0.5) -> RuntimeExpression:
    """``low`` and ``high`` alternating with ``period``, ``high`` for the
    fraction ``duty`` of each period (the ``variable_on_off_duty`` helper)."""
    on = smp.Function("variable_on_off_duty")
    return _combine(lambda lo, hi, xv, dv, per, ph: lo + (hi - lo) * on(xv, dv, 1.0 / per, ph),
                    low, high, x, duty, period, phase)


def _table_lookup(helper: str, caller: str, x, grid, values) -> RuntimeExpression:
    if not (isinstance(grid, Parameter) and isinstance(values, Parameter)):
        raise TypeError(f"{caller} requires grid and values to be registered tuple Parameters")
    return _call_function(helper, x, grid, values)


def tabulated(x, grid: Parameter, values: Parameter) -> RuntimeExpression:
    """Linear interpolation in a table held by two tuple parameters."""
    return _table_lookup("linear_interp", "tabulated", x, grid, values)


def pchip_tabulated(x, grid: Parameter, values: Parameter) -> RuntimeExpression:
    """Monotone cubic (PCHIP) interpolation in a table held by two tuple parameters."""
    return _table_lookup("pchip_interp", "pchip_tabulated", x, grid, values)


# ---- the parameter set


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


def _compound_form(expression: Any) -> smp.Expr:
    """The sympy form of a compound's right-hand side: expression text is
    parsed (helpers as function calls), numbers become constants."""
    match expression:
        case str():
            return _parse_expr(expression)
        case RuntimeExpression():
            return expression.expr
        case smp.Expr():
            return expression
        case numbers.Number():
            return smp.sympify(expression)
    raise TypeError(f"unsupported compound expression type {type(expression).__name__}")


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
        expr = _compound_form(expression)
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
    """A LindbladParameters from one, or from a mapping: names to numbers and
    expression text (the keyword form), or symbols to expressions (bindings)."""
    if isinstance(parameters, LindbladParameters):
        return parameters
    if not isinstance(parameters, Mapping):
        raise TypeError("parameters must be a LindbladParameters instance or a mapping")
    keyword_form = all(isinstance(key, str) for key in parameters) and not _is_binding_mapping(parameters)
    return LindbladParameters.from_kwargs(**dict(parameters)) if keyword_form else LindbladParameters(parameters)


def _transition_defaults(idx: int, selector: Any, gamma: float):
    """(name, value) pairs of one transition's default parameters: its Rabi
    rate (gamma, or gamma times a phase modulation), a zero detuning, and its
    polarisation (a constant 1 for one symbol; two symbols switched by a
    square wave of angular frequency gamma)."""
    rabi = str(selector.Ω)
    if getattr(selector, "phase_modulation", False):
        yield rabi, f"Ωt{idx}*phase_modulation(t, β{idx}, ωphase{idx})"
        yield f"Ωt{idx}", gamma
        yield f"β{idx}", 3.8
        yield f"ωphase{idx}", gamma
    else:
        yield rabi, gamma
    yield str(selector.δ), 0.0
    polarization = [str(sym) for sym in getattr(selector, "polarization_symbols", [])]
    if len(polarization) > 2:
        raise ValueError("polarization switching with more than two symbols is not supported")
    if len(polarization) == 1:
        yield polarization[0], 1.0
    elif len(polarization) == 2:
        first, second = polarization
        yield f"ω{idx}", gamma
        yield f"φ{idx}", 0.0
        yield first, f"square_wave(t, ω{idx}, φ{idx})"
        yield second, f"1 - {first}"


def generate_lindblad_parameters(transition_selectors: Sequence[Any], gamma: float = GAMMA_DEFAULT,
                                 **kwargs: Any) -> LindbladParameters:
    """Default parameter set for transition selectors (``parameters.py:1051-1089``);
    keyword arguments replace or add entries."""
    entries = OrderedDict(pair for idx, selector in enumerate(transition_selectors)
                          for pair in _transition_defaults(idx, selector, gamma))
    entries.update(kwargs)
    return adapt_lindblad_parameters(entries)
