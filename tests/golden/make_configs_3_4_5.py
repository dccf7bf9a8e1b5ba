"""Golden vectors of the three other benchmark workloads, from the oracle's TTGT
contraction of the same plans the GPU runs (their statevectors, 2^36 and 2^49
entries, are out of reach on one host).

Writes, under `tests/golden/`:
* `rqc_6x6_d12_s6_oracle.npz` -- 6x6 grid, depth 12, 2^6 sliced bonds: the first 16
  of the workload's 512 bitstrings and their amplitudes (every slice summed);
* `rqc_7x7_d12_s10_open5_oracle.npz` -- 7x7 grid, depth 12, 2^10 sliced bonds, qubits
  45-49 left open: the first 2 of the 32 bitstring prefixes and their 2 x 32
  amplitudes (the open legs flattened column-major, as the GPU returns them);
* `param_6x6_d6_s4_oracle.npz` -- the parameterised 6x6, depth-6 circuit with 2^4
  sliced bonds, planned as the gradient benchmark plans it: all 64 bitstrings.
Each file also holds the SHA-256 of the plan's DSL text, so a test knows it runs the
plan the vectors were made from.  Before anything is written, two amplitudes of each
plan are recomputed with the slices taken in reverse order and must agree to 1e-12.

    python tests/golden/make_configs_3_4_5.py [name ...]     # default: all three
"""

from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "qxtools.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import qxoracle as O  # noqa: E402
from qxhip import workloads  # noqa: E402
from qxhip.compute_graph import build_compute_graph  # noqa: E402
from qxhip.planning import contraction_scheme  # noqa: E402
from qxhip.tn import gate_tensor_map  # noqa: E402


def param_plan():
    """The parameterised workload's DSL and tensors, as `gradients.ParamCircuit`
    (planner "greedy") builds them."""
    w = workloads.PARAM_WORKLOAD
    tnc, _ = gate_tensor_map(w.circuit())
    tnc.add_input()
    tnc.add_output()
    groups, plan, _ = contraction_scheme(tnc, w.sliced, **w.plan_kwargs)
    g = build_compute_graph(tnc, plan, groups, dedupe=False)
    return g.to_dsl(), g.data, w.bitstrings()


def oracle_amps(dsl, data, bits):
    t = time.time()
    out = [np.asarray(O.amplitudes(dsl, data, [b]))[0].reshape(-1, order="F") for b in bits]
    print(f"  {len(bits)} bitstring(s) in {time.time() - t:.1f} s", flush=True)
    return np.stack(out)


def reversed_slices(dsl, data, bits):
    """The same amplitudes with the slice assignments summed in reverse order."""
    assigns = list(O.slice_assignments(O.parse(dsl)))[::-1]
    return np.stack([np.asarray(O.amplitudes(dsl, data, [b], slices=assigns))[0].reshape(-1, order="F")
                     for b in bits])


def main() -> None:
    jobs = [
        ("rqc_6x6_d12_s6", lambda: workloads.WORKLOADS["rqc_6x6_d12_s6"].load(), 16),
        ("rqc_7x7_d12_s10_open5", lambda: workloads.WORKLOADS["rqc_7x7_d12_s10_open5"].load(), 2),
        ("param_6x6_d6_s4", param_plan, 64),
    ]
    only = sys.argv[1:]
    for name, load, n in jobs:
        if only and name not in only:
            continue
        dsl, data, bits = load()
        bits = list(bits)[:n]
        print(name, flush=True)
        amps = oracle_amps(dsl, data, bits)
        chk = reversed_slices(dsl, data, bits[:1])
        d = np.abs(chk - amps[:1]) / np.abs(amps[:1])
        assert d.max() <= 1e-12, (name, d.max())
        sha = hashlib.sha256(dsl.encode()).hexdigest()
        np.savez_compressed(os.path.join(HERE, f"{name}_oracle.npz"), bits=np.array(bits), amps=amps,
                            dsl_sha256=np.array(sha))
        print(f"  wrote {name}_oracle.npz: {amps.shape}, reverse-slice check {d.max():.1e}", flush=True)


if __name__ == "__main__":
    main()
