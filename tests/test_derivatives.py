"""Parameter-derivative programs of the coefficient bytecode (host side of the
adjoint), checked against central differences of the original programs run
by the host expression machine on the benchmark systems."""

from __future__ import annotations

import numpy as np
import pytest

from centrex_tlf_amd.lindblad import prepare_lindblad_problem
from centrex_tlf_amd.lindblad.derivatives import differentiate_flat_plan
from centrex_tlf_amd.lindblad.ir import CompiledExpression, evaluate_expression

from .systems import GAMMA, tlf_n39, tlf_n39_xy, tlf_n72


def _programs(flat):
    rows = list(zip(flat["instr_op"], flat["instr_index"], flat["instr_argc"], flat["instr_function"],
                    flat["instr_re"], flat["instr_im"]))
    ptrs = flat["expr_ptrs"]
    return [CompiledExpression(tuple(tuple(r) for r in rows[ptrs[e]:ptrs[e + 1]])) for e in range(len(ptrs) - 1)]


def _evaluate(flat, base, t):
    """Slots and all coefficient programs (incl. derivative programs) at time t."""
    progs = _programs(flat)
    slots = [complex(v) for v in base] + [0j] * (int(flat["n_slots"]) - len(base))
    nc = int(flat["n_compounds"])
    for c in range(nc):
        slots[int(flat["compound_slots"][c])] = evaluate_expression(progs[c], slots, t)
    return np.asarray([evaluate_expression(p, slots, t) for p in progs[nc:]], dtype=np.complex128)


@pytest.mark.parametrize("builder,names", [
    (tlf_n39_xy, ["delta0", "Omega0", "theta", "sigma"]),
    (tlf_n39, ["delta0", "Omega0", "sigma", "t_center"]),
    (tlf_n72, ["delta0", "Omega0", "sigma"]),
])
def test_derivative_programs_match_central_differences(builder, names):
    system, params, _, t_span, _ = builder()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    flat = prepared.flat_plan
    slots = [prepared.slot_names.index(n) for n in names]
    dflat = differentiate_flat_plan(flat, slots)
    K, P = int(flat["n_coefficients"]), len(slots)
    base = np.asarray(flat["base_re"]) + 1j * np.asarray(flat["base_im"])
    rng = np.random.default_rng(5)
    for t in rng.uniform(*t_span, size=4):
        got = _evaluate(dflat, base, t)
        assert got.size == K + K * P
        np.testing.assert_allclose(got[:K], _evaluate(flat, base, t), rtol=1e-15, atol=0)
        for m, s in enumerate(slots):
            hstep = 1e-6 * abs(base[s]) if base[s] != 0 else 1e-6 * GAMMA
            hi, lo = base.copy(), base.copy()
            hi[s] += hstep
            lo[s] -= hstep
            fd = (_evaluate(flat, hi, t) - _evaluate(flat, lo, t)) / (2 * hstep)
            an = got[K + np.arange(K) * P + m]
            scale = np.abs(fd).max() + np.abs(an).max() + 1e-300
            np.testing.assert_allclose(an, fd, rtol=0, atol=1e-6 * scale, err_msg=f"{names[m]} t={t}")


def test_derivative_rules_reject_unsupported():
    system, params, _, _, _ = tlf_n39()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    with pytest.raises(ValueError, match="base parameter"):
        differentiate_flat_plan(prepared.flat_plan, [int(prepared.flat_plan["n_slots"]) - 1])


@pytest.mark.parametrize("builder", [tlf_n39_xy, tlf_n39, tlf_n72])
def test_time_derivative_programs_match_central_differences(builder):
    """The time column (with_time=True): d c_k / d t against central differences
    in t, with the parameter columns unchanged."""
    system, params, _, t_span, _ = builder()
    prepared = prepare_lindblad_problem(system, params, backend="host")
    flat = prepared.flat_plan
    slots = [prepared.slot_names.index(n) for n in ("delta0", "Omega0")]
    dflat = differentiate_flat_plan(flat, slots, with_time=True)
    plain = differentiate_flat_plan(flat, slots)
    assert dflat["diff_time"] == 1 and plain["diff_time"] == 0
    K, P = int(flat["n_coefficients"]), len(slots)
    base = np.asarray(flat["base_re"]) + 1j * np.asarray(flat["base_im"])
    rng = np.random.default_rng(9)
    for t in rng.uniform(*t_span, size=6):
        got = _evaluate(dflat, base, t)
        assert got.size == K + K * (P + 1)
        ref = _evaluate(plain, base, t)
        for m in range(P):
            np.testing.assert_array_equal(got[K + np.arange(K) * (P + 1) + m], ref[K + np.arange(K) * P + m])
        ht = 1e-6 * (t_span[1] - t_span[0])
        fd = (_evaluate(flat, base, t + ht) - _evaluate(flat, base, t - ht)) / (2 * ht)
        an = got[K + np.arange(K) * (P + 1) + P]
        scale = np.abs(fd).max() + np.abs(an).max() + 1e-300
        np.testing.assert_allclose(an, fd, rtol=0, atol=1e-6 * scale, err_msg=f"t={t}")
