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
                                                    output_args=out, **self.plan_kwargs)

    def load(self) -> tuple:
        """(dsl, data, bitstrings) of the saved plan."""
        dsl, data, params = simulation.load_simulation_files(self.prefix())
        bits = params.get("output", {}).get("params", {}).get("bitstrings") or self.bitstrings()
        return dsl, data, list(bits)


WORKLOADS = {w.name: w for w in [
    # plain method, small: 3x3 grid, depth 16, 64 bitstrings, complex128, no slicing
    Workload("rqc_3x3_d16", 3, 3, 16, 64, "c128"),
    # plain method at scale: 5x5 grid, depth 20, 256 bitstrings, complex128, no slicing
    Workload("rqc_5x5_d20", 5, 5, 20, 256, "c128"),
]}


def main(argv):
    names = argv or list(WORKLOADS)
    for n in names:
        w = WORKLOADS[n]
        g, tnc, plan, groups = w.generate()
        print(f"{n}: {len(tnc)} tensors, {len(plan)} steps, {len(groups)} sliced -> {w.prefix()}.qx")


if __name__ == "__main__":
    main(sys.argv[1:])
