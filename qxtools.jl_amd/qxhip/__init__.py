"""qxhip: MI355X (gfx950) executor for QXTools' tensor-network contraction path.

Public names follow the reference module (`src/QXTools.jl:1-25`): circuits,
`convert_to_tnc`, `TensorNetworkCircuit`, planning (`contraction_scheme`),
`build_compute_graph`, `TensorCache`, and the simulation API (`single_amplitude`,
`run_simulation`, `amplitudes_all`, `amplitudes_uniform`,
`generate_simulation_files`).  Contractions run in libqxhip.so (HIP, gfx950).
"""

from .circuits import (Circuit, Gate, create_ghz_circuit, create_param_rqc_circuit,  # noqa: F401
                       create_rqc_circuit, create_test_circuit, gate_matrix, gate_qubits)
from .compute_graph import ComputeGraph, TensorCache, build_compute_graph, parse_dsl  # noqa: F401
from .planning import (contraction_plan, contraction_scheme, greedy_contraction_plan,  # noqa: F401
                       order_to_contraction_plan, sweep_contraction_plan)
from .simulation import (Executor, amplitudes_all, amplitudes_uniform,  # noqa: F401
                         generate_simulation_files, load_simulation_files, output_params_dict,
                         run_outputs, run_simulation, single_amplitude)
from .tn import TensorNetwork, TensorNetworkCircuit, convert_to_tnc  # noqa: F401
