"""The five benchmark workloads and their saved plans.

Each workload is a seeded grid circuit, a contraction plan computed once and saved
as `<name>.qx` + `<name>.npz` (+ `.yml` with the bitstrings) under
`qxtools.jl_amd/workloads/`, and a seeded bitstring list -- the role of
`generate_simulation_files` (`src/simulation.jl:52-77`) and
`bin/prepare_rqc_simulation_files.jl`.

    python -m qxhip.workloads [name ...]     # (re)write saved plans
"""

from __future__ import annotations

import os
import sys
from dataclasses import dataclass, field

from . import circuits, simulation

WORKLOAD_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "workloads")

CIRCUIT_SEED = 42      # docs/src/distributed.md:15 uses seed 42 for its RQC
BITSTRING_SEED = 4242
PARAM_SEED = 4243


@dataclass
class Workload:
    name: str
    rows: int
    cols: int
    depth: int
    num_bitstrings: int
    precision: str
    sliced: int = 0
    parameterised: bool = False
    plan_kwargs: dict = field(default_factory=dict)
    open_outputs: tuple = ()          # qubits left open: 2^m amplitudes per bitstring
    note: str = ""
    batch_last: int = 0               # executor layout option (batched leg slowest)

    @property
    def qubits(self) -> int:
        return self.rows * self.cols

    def circuit(self) -> circuits.Circuit:
        if self.parameterised:
            return circuits.create_param_rqc_circuit(self.rows, self.cols, self.depth, seed=PARAM_SEED)
        return circuits.create_rqc_circuit(self.rows, self.cols, self.depth, seed=CIRCUIT_SEED)

    def bitstrings(self) -> list:
        return simulation.amplitudes_uniform(self.qubits, BITSTRING_SEED, self.num_bitstrings)

    def prefix(self) -> str:
        return os.path.join(WORKLOAD_DIR, self.name)

    def generate(self):
        os.makedirs(WORKLOAD_DIR, exist_ok=True)
        out = simulation.output_params_dict(self.qubits, self.num_bitstrings, "List",
                                            bitstrings=self.bitstrings())
        return simulation.generate_simulation_files(self.circuit(), self.prefix(), self.sliced,
                                                    output_args=out, open_outputs=self.open_outputs,
                                                    **self.plan_kwargs)

    def load(self) -> tuple:
        """(dsl, data, bitstrings) of the saved plan."""
        dsl, data, params = simulation.load_simulation_files(self.prefix())
        bits = params.get("output", {}).get("params", {}).get("bitstrings") or self.bitstrings()
        return dsl, data, list(bits)


WORKLOADS = {w.name: w for w in [
    # plain method, small: 3x3 grid, depth 16, 64 bitstrings, complex128, no slicing
    Workload("rqc_3x3_d16", 3, 3, 16, 64, "c128"),
    # plain method at scale: 5x5 grid, depth 20, 256 bitstrings, complex128, no slicing
    Workload("rqc_5x5_d20", 5, 5, 20, 256, "c128", plan_kwargs={"max_block_qubits": 4}),
    # precision: 6x6 grid, complex64 storage, 2^6 slices, 512 bitstrings.  Depth 12
    # instead of 24: at depth 24 the planners here reach width 39 (2^58 flops per
    # amplitude), beyond a single-GPU benchmark.
    Workload("rqc_6x6_d12_s6", 6, 6, 12, 512, "c64_4m", sliced=6,
             plan_kwargs={"trials": 16, "min_fill_trials": 4},
             note="depth reduced from 24 to 12", batch_last=1),
    # batching: 7x7 grid, 2^10 slices, 5 output legs left open so each of 32 bitstring
    # prefixes yields 32 amplitudes (1024 in all).  Depth 12 instead of 24 (width).
    Workload("rqc_7x7_d12_s10_open5", 7, 7, 12, 32, "c64_4m", sliced=10,
             plan_kwargs={"trials": 16, "min_fill_trials": 4}, open_outputs=(45, 46, 47, 48, 49),
             note="depth reduced from 24 to 12", batch_last=1),
]}

# differentiation: parameterised 6x6 grid, Rz Ry Rz per qubit per cycle, 2^4 slices,
# 64 bitstrings.  Depth 6 instead of 20 (648 angles); built at run time by
# qxhip.gradients.ParamCircuit, which needs one data key per gate.
PARAM_WORKLOAD = Workload("param_6x6_d6_s4", 6, 6, 6, 64, "c128", sliced=4, parameterised=True,
                          plan_kwargs={"trials": 16, "min_fill_trials": 4},
                          note="depth reduced from 20 to 6")


def main(argv):
    names = argv or list(WORKLOADS)
    for n in names:
        w = WORKLOADS[n]
        g, tnc, plan, groups = w.generate()
        print(f"{n}: {len(tnc)} tensors, {len(plan)} steps, {len(groups)} sliced -> {w.prefix()}.qx")


if __name__ == "__main__":
    main(sys.argv[1:])
