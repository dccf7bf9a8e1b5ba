"""Golden vectors of the 5x5-grid, depth-20 workload (25 qubits, complex128, no slicing).

Writes `tests/golden/rqc_5x5_d20_amplitudes.npz`:
* `bits`        -- the workload's 256 seeded bitstrings (as stored in its saved plan);
* `sv_amps`     -- their amplitudes from the oracle's FP64 statevector simulator
                   (2^25 complex128 = 512 MiB), which shares no code with the
                   tensor-network path;
* `oracle_idx`, `oracle_amps` -- 8 of them contracted over the saved plan by the
                   oracle's TTGT executor (the same plan file the GPU runs).

The two references are checked against each other here before anything is written.

    python tests/golden/make_rqc_5x5_d20.py
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "qxtools.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import qxoracle as O  # noqa: E402
from qxhip import circuits, workloads  # noqa: E402

OUT = os.path.join(HERE, "rqc_5x5_d20_amplitudes.npz")
ORACLE_PICK = 8


def main() -> None:
    w = workloads.WORKLOADS["rqc_5x5_d20"]
    dsl, data, bits = w.load()
    assert bits == w.bitstrings(), "saved plan's bitstrings differ from the seeded generator"
    circ = w.circuit()
    gates = [(circuits.gate_matrix(g), circuits.gate_qubits(g)) for g in circ.ops]
    t0 = time.time()
    psi = O.statevector(circ.num_qubits, gates)
    print(f"statevector: {len(gates)} gates in {time.time() - t0:.1f} s", flush=True)
    # all characters are '0'/'1': an amplitude is one entry (axis q-1 = qubit q)
    sv = np.array([psi[tuple(int(c) for c in b)] for b in bits])
    norm = float(np.vdot(psi.reshape(-1), psi.reshape(-1)).real)
    assert abs(norm - 1.0) < 1e-12, norm
    del psi
    idx = np.random.Generator(np.random.PCG64(5)).choice(len(bits), ORACLE_PICK, replace=False)
    idx.sort()
    t0 = time.time()
    orc = O.amplitudes(dsl, data, [bits[i] for i in idx])
    print(f"oracle: {ORACLE_PICK} amplitudes in {time.time() - t0:.1f} s", flush=True)
    rel = np.abs(orc - sv[idx]) / np.abs(sv[idx])
    print(f"oracle vs statevector: max rel err {rel.max():.3e}")
    assert rel.max() < 1e-11
    np.savez(OUT, bits=np.array(bits), sv_amps=sv, oracle_idx=idx, oracle_amps=orc)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
