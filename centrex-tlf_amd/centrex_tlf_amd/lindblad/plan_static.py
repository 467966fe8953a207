"""``prepare_lindblad_problem``: lower an OBE system once, upload it to the GPU once.

Follows ``centrex_tlf/lindblad/plan_static.py:137-185``: parameters are adapted,
the parameter graph and the decomposed Hamiltonian are lowered, the collapse
operators are reduced to single jumps, and the expanded sparse RHS program is
built.  Instead of a PyO3 plan the result carries a device plan created through
the C ABI (``lb_plan_create``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .ir import (flatten_plan, lower_expanded_sparse_rhs, lower_hamiltonian_decomposed, lower_hamiltonian_entrywise,
                 lower_parameter_graph)
from .parameters import LindbladParameters, adapt_lindblad_parameters
from .state_layout import PackedHermitianLayout

__all__ = ["PreparedLindbladProblem", "prepare_lindblad_problem", "lower_structured_jumps"]

# Backends whose prepared problem owns a device plan.  "rust" is accepted so
# code written for the reference runs unchanged and lands on the GPU.
_DEVICE_BACKENDS = {"gpu", "rust"}
_HOST_BACKENDS = {"host"}


@dataclass
class PreparedLindbladProblem:
    obe_system: Any
    parameters: LindbladParameters
    layout: PackedHermitianLayout
    parameter_graph: dict[str, Any]
    hamiltonian_plan: dict[str, Any]
    dense_c_array: np.ndarray
    structured_jumps: list[dict[str, Any]]
    source_decay_rates: np.ndarray
    incoming_transfers_by_target: list[list[dict[str, Any]]]
    expanded_rhs_plan: dict[str, Any] | None
    flat_plan: dict[str, Any]
    backend: str
    gpu_plan: Any | None = field(default=None, repr=False)
    # Adaptive solves run kernels with the time-dependent coefficient programs
    # compiled in (DevicePlan.specialize); False keeps the precompiled kernels
    # with the expression machine.  Results are bitwise the same.
    specialize: bool = True

    def check_execution_mode(self, execution_mode: str) -> None:
        """The reference's fail-fast check (``plan_static.py:37-56``): its
        ``expanded_sparse`` modes need a decomposed plan.  (The GPU runs every
        plan through its expanded sparse program; the check keeps the
        reference's error for the same request.)"""
        if "expanded_sparse" not in execution_mode:
            return
        kind = self.hamiltonian_plan.get("kind", "entrywise")
        if kind == "decomposed":
            return
        raise ValueError(
            f"execution_mode={execution_mode!r} requires a decomposed Hamiltonian "
            f"plan, but this problem was prepared with kind={kind!r}. Either pass "
            "hamiltonian_representation='decomposed' to prepare_lindblad_problem, "
            "or use execution_mode='structured'."
        )

    def device_plan(self):
        """The device plan, created on first use."""
        if self.gpu_plan is None:
            from .native import DevicePlan

            self.gpu_plan = DevicePlan(self.flat_plan)
            if self.specialize:
                from .native import NativeError

                try:
                    self.gpu_plan.specialize()
                except NativeError as exc:  # an instruction without a generated form
                    import warnings

                    warnings.warn(f"plan not specialised, precompiled kernels used: {exc}", stacklevel=2)
        return self.gpu_plan

    @property
    def slot_names(self) -> list[str]:
        return list(self.parameter_graph["slot_names"])

    @property
    def base_count(self) -> int:
        return len(self.parameter_graph["base_values"])


def lower_structured_jumps(c_array: np.ndarray):
    """Single-jump collapse operators -> rates (``plan_static.py:107-134``).

    Each ``C_k`` must have exactly one nonzero ``C_k[target, source]``; its rate
    ``|C|^2`` leaves ``source`` and arrives at ``target``.
    """
    ops = np.asarray(c_array, dtype=np.complex128)
    k, n = int(ops.shape[0]), int(ops.shape[1])
    present = (np.abs(ops) > 0).reshape(k, n * n)
    counts = present.sum(axis=1)
    wrong = np.flatnonzero(counts != 1)
    if wrong.size:
        first = int(wrong[0])
        raise AssertionError(
            f"collapse operator {first} is not a single-jump operator; found {int(counts[first])} nonzero elements")
    target, source = np.divmod(present.argmax(axis=1), n)
    amp = ops[np.arange(k), target, source]
    rate = (amp * amp.conj()).real
    # bincount adds in operator order, as a running sum over the operators would
    source_decay = np.bincount(source, weights=rate, minlength=n).astype(np.float64)
    jumps = [{"target": int(t), "source": int(s_), "amp_re": float(a.real), "amp_im": float(a.imag),
              "rate": float(r)} for t, s_, a, r in zip(target, source, amp, rate)]
    incoming: list[list[dict[str, Any]]] = [[] for _ in range(n)]
    for jump in jumps:
        incoming[jump["target"]].append({"source": jump["source"], "rate": jump["rate"]})
    return jumps, source_decay, incoming


def prepare_lindblad_problem(obe_system: Any, parameters: Any, backend: str = "gpu",
                             hamiltonian_representation: str = "decomposed") -> PreparedLindbladProblem:
    """Lower ``obe_system`` with ``parameters`` and (for a device backend) upload it.

    ``backend="gpu"`` (alias ``"rust"``) creates the device plan immediately;
    ``backend="host"`` only lowers, which is enough to inspect the tables.
    ``hamiltonian_representation``: "decomposed" (the default, also used for
    "auto") or "entrywise" (one coefficient per distinct upper entry).
    """
    if backend not in _DEVICE_BACKENDS | _HOST_BACKENDS:
        raise ValueError(f"unsupported backend {backend!r}; use 'gpu' (or 'rust') or 'host'")
    if hamiltonian_representation not in {"auto", "entrywise", "decomposed"}:
        raise ValueError("representation must be one of 'auto', 'entrywise', or 'decomposed'")
    params = adapt_lindblad_parameters(parameters)
    params.check_hamiltonian_symbols(obe_system.H_symbolic)
    n = int(obe_system.H_symbolic.shape[0])
    layout = PackedHermitianLayout(n)
    graph = lower_parameter_graph(params)
    lower = lower_hamiltonian_entrywise if hamiltonian_representation == "entrywise" else lower_hamiltonian_decomposed
    hamiltonian_plan = lower(obe_system.H_symbolic, params.slot_index_by_name, graph["tuple_value_names"])
    dense_c = np.asarray(obe_system.C_array, dtype=np.complex128)
    jumps, source_decay, incoming = lower_structured_jumps(dense_c)
    rhs_plan = lower_expanded_sparse_rhs(hamiltonian_plan, source_decay, incoming)
    flat = flatten_plan(n, graph, hamiltonian_plan, rhs_plan)
    prepared = PreparedLindbladProblem(
        obe_system=obe_system,
        parameters=params,
        layout=layout,
        parameter_graph=graph,
        hamiltonian_plan=hamiltonian_plan,
        dense_c_array=dense_c,
        structured_jumps=jumps,
        source_decay_rates=source_decay,
        incoming_transfers_by_target=incoming,
        expanded_rhs_plan=rhs_plan,
        flat_plan=flat,
        backend=backend,
    )
    if backend in _DEVICE_BACKENDS:
        prepared.device_plan()
    return prepared
