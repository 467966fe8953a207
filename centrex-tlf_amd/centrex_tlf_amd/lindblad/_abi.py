"""ctypes mirror of ``include/centrex_lindblad.h``.

Field order and types must match the header exactly; ``tests/test_abi.py``
checks the struct sizes against the C compiler's view.
"""

from __future__ import annotations

import ctypes as C
from typing import Any

import numpy as np

__all__ = [
    "LB_OK",
    "LB_MODE_PERSISTENT",
    "LB_MODE_PLAIN",
    "LB_ADJOINT",
    "LB_EVENT",
    "LB_OUT",
    "LB_TAPE",
    "LB_SOLVER",
    "LB_TRAJ",
    "LbPlanDesc",
    "LbEffPlanDesc",
    "LbEffSolveArgs",
    "LbSolveArgs",
    "make_plan_desc",
    "ptr",
]

LB_OK = 0
LB_SOLVER = {"dopri5": 0, "tsit5": 1, "fixed_dopri5": 2, "fixed_rk2": 3, "fixed_rk4": 4}
LB_OUT = {"populations": 0, "selected": 1, "integral": 2, "full": 3, "rate": 4}
LB_MODE_PERSISTENT = 0
LB_MODE_PLAIN = 1
LB_TRAJ = {0: "ok", 1: "maxiters", 2: "underflow", 3: "stiff", 4: "bad_span", 5: "bad_support",
           6: "saveat", 7: "capacity", 8: "bad_dt", 9: "event_complex", 10: "field_below", 11: "field_above"}
LB_EVENT = {"none": 0, "population": 1, "runtime_expression": 2}
# attempt-tape flags and adjoint modes
LB_TAPE = {"accepted": 1, "clipped": 2, "clamped": 4, "hmax": 8, "minrule": 16, "errq": 32, "first": 64}
LB_ADJOINT = {"adaptive": 0, "frozen": 1}

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


class LbPlanDesc(C.Structure):
    _fields_ = [
        ("n_states", C.c_int32),
        ("n_slots", C.c_int32),
        ("n_base", C.c_int32),
        ("base_re", _f64p),
        ("base_im", _f64p),
        ("n_compounds", C.c_int32),
        ("compound_slots", _i32p),
        ("n_coefficients", C.c_int32),
        ("expr_ptrs", _i32p),
        ("instr_op", _i32p),
        ("instr_index", _i32p),
        ("instr_argc", _i32p),
        ("instr_function", _i32p),
        ("instr_re", _f64p),
        ("instr_im", _f64p),
        ("n_terms", C.c_int64),
        ("output_ptrs", _i64p),
        ("input_indices", _i64p),
        ("coeff_indices", _i64p),
        ("coeff_conj", _u8p),
        ("input_conj", _u8p),
        ("factors_re", _f64p),
        ("factors_im", _f64p),
        ("n_diff_params", C.c_int32),
        ("diff_param_slots", _i32p),
        ("diff_time", C.c_int32),
    ]


class LbSolveArgs(C.Structure):
    _fields_ = [
        ("n_traj", C.c_int64),
        ("solver", C.c_int32),
        ("t0", C.c_double),
        ("t1", C.c_double),
        ("t1_per_traj", C.c_void_p),
        ("abstol", C.c_double),
        ("reltol", C.c_double),
        ("dt", C.c_double),
        ("maxiters", C.c_int64),
        ("save_start", C.c_int32),
        ("y0", C.c_void_p),
        ("y0_shared", C.c_int32),
        ("n_override", C.c_int32),
        ("override_slots", _i32p),
        ("override_values", C.c_void_p),
        ("n_axes", C.c_int32),
        ("axis_lengths", _i64p),
        ("axis_values", C.c_void_p),
        ("output_kind", C.c_int32),
        ("n_selected", C.c_int32),
        ("selected_rows", _i32p),
        ("selected_cols", _i32p),
        ("n_weights", C.c_int32),
        ("weight_indices", _i32p),
        ("weight_values", _f64p),
        ("out_values", C.c_void_p),
        ("out_stats", C.c_void_p),
        ("out_fail_t", C.c_void_p),
        ("mode", C.c_int32),
        ("threads_per_traj", C.c_int32),
        ("workgroups", C.c_int32),
        ("out_integral", C.c_void_p),
        ("variant", C.c_int32),
        ("n_saveat", C.c_int32),
        ("saveat", C.c_void_p),
        ("trace", C.c_int32),
        ("capacity", C.c_int64),
        ("out_times", C.c_void_p),
        ("out_counts", C.c_void_p),
        ("y0_support", C.POINTER(C.c_uint8)),
        ("tape_t", C.c_void_p),
        ("tape_h", C.c_void_p),
        ("tape_y", C.c_void_p),
        ("tape_n", C.c_void_p),
        ("tape_cap", C.c_int64),
        ("ckpt_every", C.c_int32),
        ("out_grad", C.c_void_p),
        ("replay_h", C.c_void_p),
        ("replay_n", C.c_void_p),
        ("queue_order", C.c_void_p),
        ("event_kind", C.c_int32),
        ("n_event_indices", C.c_int32),
        ("event_indices", _i32p),
        ("event_threshold", C.c_double),
        ("n_event_instr", C.c_int32),
        ("event_op", _i32p),
        ("event_index", _i32p),
        ("event_argc", _i32p),
        ("event_function", _i32p),
        ("event_re", _f64p),
        ("event_im", _f64p),
        ("out_event_t", C.c_void_p),
        ("tape_c", C.c_void_p),
        ("tape_f", C.c_void_p),
        ("adjoint_mode", C.c_int32),
    ]


class LbEffPlanDesc(C.Structure):
    _fields_ = [
        ("n_states", C.c_int32),
        ("real_dim", C.c_int32),
        ("n_grid", C.c_int32),
        ("field_grid", _f64p),
        ("row_ptrs", C.c_void_p * 3),
        ("col_indices", C.c_void_p * 3),
        ("coeffs", C.c_void_p * 3),
        ("nnz", C.c_int64 * 3),
        ("field_slot", C.c_int32),
        ("rabi_slot", C.c_int32),
        ("detuning_slot", C.c_int32),
        ("n_slots", C.c_int32),
        ("n_base", C.c_int32),
        ("base_re", _f64p),
        ("base_im", _f64p),
        ("n_compounds", C.c_int32),
        ("compound_slots", _i32p),
        ("expr_ptrs", _i32p),
        ("instr_op", _i32p),
        ("instr_index", _i32p),
        ("instr_argc", _i32p),
        ("instr_function", _i32p),
        ("instr_re", _f64p),
        ("instr_im", _f64p),
    ]


class LbEffSolveArgs(C.Structure):
    _fields_ = [
        ("n_traj", C.c_int64),
        ("solver", C.c_int32),
        ("t0", C.c_double),
        ("t1", C.c_double),
        ("abstol", C.c_double),
        ("reltol", C.c_double),
        ("dt", C.c_double),
        ("maxiters", C.c_int64),
        ("save_start", C.c_int32),
        ("y0", C.c_void_p),
        ("n_override", C.c_int32),
        ("override_slots", _i32p),
        ("override_values", C.c_void_p),
        ("output_kind", C.c_int32),
        ("n_selected", C.c_int32),
        ("selected_rows", _i32p),
        ("selected_cols", _i32p),
        ("n_weights", C.c_int32),
        ("weight_indices", _i32p),
        ("weight_values", _f64p),
        ("n_saveat", C.c_int32),
        ("saveat", C.c_void_p),
        ("capacity", C.c_int64),
        ("out_values", C.c_void_p),
        ("out_times", C.c_void_p),
        ("out_counts", C.c_void_p),
        ("out_stats", C.c_void_p),
        ("out_fail_t", C.c_void_p),
        ("workgroups", C.c_int32),
    ]


LB_EFF_OUT = {"populations": 0, "selected": 1, "integral": 2, "full": 3}


def ptr(array: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (NULL for None)."""
    if array is None:
        return C.cast(None, C.POINTER(ctype))
    assert array.flags["C_CONTIGUOUS"]
    return array.ctypes.data_as(C.POINTER(ctype))


def make_plan_desc(flat: dict[str, Any]) -> tuple[LbPlanDesc, list[np.ndarray]]:
    """Build an ``lb_plan_desc`` from :func:`ir.flatten_plan` output.

    Returns the struct plus the arrays it points into (keep them alive for
    the duration of the create call).
    """
    keep: list[np.ndarray] = []

    def arr(key: str, dtype) -> np.ndarray:
        a = np.ascontiguousarray(flat[key], dtype=dtype)
        if a.size == 0:
            a = np.zeros(1, dtype=dtype)
        keep.append(a)
        return a

    d = LbPlanDesc()
    d.n_states = int(flat["n_states"])
    d.n_slots = int(flat["n_slots"])
    d.n_base = int(flat["n_base"])
    d.base_re = ptr(arr("base_re", np.float64), C.c_double)
    d.base_im = ptr(arr("base_im", np.float64), C.c_double)
    d.n_compounds = int(flat["n_compounds"])
    d.compound_slots = ptr(arr("compound_slots", np.int32), C.c_int32)
    d.n_coefficients = int(flat["n_coefficients"])
    d.expr_ptrs = ptr(arr("expr_ptrs", np.int32), C.c_int32)
    d.instr_op = ptr(arr("instr_op", np.int32), C.c_int32)
    d.instr_index = ptr(arr("instr_index", np.int32), C.c_int32)
    d.instr_argc = ptr(arr("instr_argc", np.int32), C.c_int32)
    d.instr_function = ptr(arr("instr_function", np.int32), C.c_int32)
    d.instr_re = ptr(arr("instr_re", np.float64), C.c_double)
    d.instr_im = ptr(arr("instr_im", np.float64), C.c_double)
    d.n_terms = int(np.asarray(flat["input_indices"]).size)
    d.output_ptrs = ptr(arr("output_ptrs", np.int64), C.c_int64)
    d.input_indices = ptr(arr("input_indices", np.int64), C.c_int64)
    d.coeff_indices = ptr(arr("coeff_indices", np.int64), C.c_int64)
    d.coeff_conj = ptr(arr("coeff_conj", np.uint8), C.c_uint8)
    d.input_conj = ptr(arr("input_conj", np.uint8), C.c_uint8)
    d.factors_re = ptr(arr("factors_re", np.float64), C.c_double)
    d.factors_im = ptr(arr("factors_im", np.float64), C.c_double)
    d.n_diff_params = int(flat.get("n_diff_params", 0))
    if d.n_diff_params:
        d.diff_param_slots = ptr(arr("diff_param_slots", np.int32), C.c_int32)
    d.diff_time = int(flat.get("diff_time", 0))
    return d, keep
