"""Write the TlF OBE systems used by the benchmark configurations as data fixtures.

The molecular builders (states, reduced Hamiltonians, couplings, branching
ratios) are not part of this project's hot path, so the two systems the
workloads are quoted on are generated once here from the reference's own
builders and committed as plain JSON:

* ``tlf_r0_f1_1o2_f1_n39.json``: X(J=0,1,2) + B(J=1, F1'=1/2, F'=1), one
  z-polarised laser on R(0) F1'=1/2 F'=1 (n = 39).
* ``tlf_r0_f1_1o2_f1_xy_n39.json``: the same n = 39 levels with the laser's
  polarisation split into x and y components (symbols PX0, PY0), the system
  the gradient workload differentiates with respect to the polarisation angle.
* ``tlf_r0_f1_3o2_n72.json``: X(J=0..3) + B(J=1, F1'=3/2, F'=1,2), two lasers
  on R(0) F1'=3/2 F'=2 and F'=1 with x/y polarisation switching, plus a
  J=1 -> J=2 microwave (n = 72).

Each file holds the nonzero symbolic Hamiltonian entries (``sympy.srepr``),
the single-jump collapse operators as ``(target, source, amplitude)`` and the
state labels.  Builder calls follow
``centrex_tlf/lindblad/utils_setup.py:413`` (``generate_OBE_system``)
and ``centrex_tlf/couplings/transition.py:110``
(``generate_transition_selectors``); the decay rate is passed explicitly as
Gamma = 2*pi*1.56 MHz.

Run from the repository root:  ``python tests/golden/make_tlf_systems.py [name ...]``
(names: n39, xy39, n72; default all).
"""

from __future__ import annotations

import json
import math
import sys
import time
from pathlib import Path

import numpy as np
import sympy as smp

REFERENCE = Path("/root/reference")
OUT_DIR = Path(__file__).resolve().parent
GAMMA = 2.0 * math.pi * 1.56e6


def _dump_system(obe, path: Path, description: str) -> None:
    H = obe.H_symbolic
    n = int(H.shape[0])
    entries = []
    for i in range(n):
        for j in range(n):
            value = H[i, j]
            if value != 0:
                entries.append([i, j, smp.srepr(value)])
    C = np.asarray(obe.C_array, dtype=np.complex128)
    jumps = []
    for k, collapse in enumerate(C):
        nz = np.argwhere(np.abs(collapse) > 0)
        for target, source in nz:
            amp = complex(collapse[target, source])
            jumps.append([int(k), int(target), int(source), amp.real, amp.imag])
    labels = []
    for state in obe.QN:
        s = state.largest
        labels.append(
            {
                "electronic": str(s.electronic_state.name),
                "J": int(s.J),
                "F1": float(s.F1),
                "F": int(s.F),
                "mF": int(s.mF),
                "P": int(s.P) if s.P is not None else 0,
            }
        )
    payload = {
        "description": description,
        "n": n,
        "gamma": GAMMA,
        "n_collapse": int(C.shape[0]),
        "hamiltonian_entries": entries,
        "collapse_jumps": jumps,
        "states": labels,
        "symbols": sorted(str(s) for s in H.free_symbols),
    }
    path.write_text(json.dumps(payload, indent=0))
    print(f"wrote {path} n={n} entries={len(entries)} jumps={len(jumps)}", flush=True)


def main() -> None:
    sys.path.insert(0, str(REFERENCE))
    from centrex_tlf import couplings, lindblad, states, transitions  # noqa: E402

    X = states.ElectronicState.X
    B = states.ElectronicState.B
    wanted = set(sys.argv[1:]) or {"n39", "xy39", "n72"}

    if "xy39" in wanted:
        t0 = time.time()
        trans = [transitions.R0_F1_1o2_F1]
        selectors = couplings.generate_transition_selectors(
            trans, [[couplings.polarization_X, couplings.polarization_Y]])
        obe = lindblad.generate_OBE_system(
            states.QuantumSelector(J=[0, 1, 2], electronic=X),
            states.QuantumSelector(J=1, F1=1 / 2, F=1, P=trans[0].P_excited, electronic=B),
            selectors,
            Γ=GAMMA,
        )
        _dump_system(
            obe,
            OUT_DIR / "tlf_r0_f1_1o2_f1_xy_n39.json",
            "X(J=0,1,2) + B(J=1,F1'=1/2,F'=1); R0_F1_1o2_F1, x and y polarisation components",
        )
        print(f"xy n39 built in {time.time() - t0:.1f} s", flush=True)
    if "n39" not in wanted and "n72" not in wanted:
        return

    t0 = time.time()
    trans = [transitions.R0_F1_1o2_F1]
    selectors = couplings.generate_transition_selectors(trans, [[couplings.polarization_Z]])
    obe = lindblad.generate_OBE_system(
        states.QuantumSelector(J=[0, 1, 2], electronic=X),
        states.QuantumSelector(J=1, F1=1 / 2, F=1, P=trans[0].P_excited, electronic=B),
        selectors,
        Γ=GAMMA,
    )
    _dump_system(
        obe,
        OUT_DIR / "tlf_r0_f1_1o2_f1_n39.json",
        "X(J=0,1,2) + B(J=1,F1'=1/2,F'=1); R0_F1_1o2_F1, z polarisation",
    )
    print(f"n39 built in {time.time() - t0:.1f} s", flush=True)

    t0 = time.time()
    trans = [
        transitions.R0_F1_3o2_F2,
        transitions.R0_F1_3o2_F1,
        transitions.MicrowaveTransition(1, 2),
    ]
    pols = [
        [couplings.polarization_X, couplings.polarization_Y],
        [couplings.polarization_X, couplings.polarization_Y],
        [couplings.polarization_Z],
    ]
    selectors = couplings.generate_transition_selectors(trans, pols)
    obe = lindblad.generate_OBE_system(
        states.QuantumSelector(J=[0, 1, 2, 3], electronic=X),
        states.QuantumSelector(J=1, F1=3 / 2, F=[1, 2], P=trans[0].P_excited, electronic=B),
        selectors,
        Γ=GAMMA,
    )
    _dump_system(
        obe,
        OUT_DIR / "tlf_r0_f1_3o2_n72.json",
        "X(J=0..3) + B(J=1,F1'=3/2,F'=1,2); R0_F1_3o2_F2 + R0_F1_3o2_F1 with x/y "
        "polarisation, microwave J=1->J=2 z polarisation",
    )
    print(f"n72 built in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
