"""Simulation API: the drop-in for `src/simulation.jl` and `src/outputs.jl`.

* `amplitudes_all`, `amplitudes_uniform`        -- `src/simulation.jl:14-28`
* `single_amplitude(tnc, plan, bitstring)`      -- `src/simulation.jl:86-91`
* `run_simulation(circ; num_amplitudes, seed)`  -- `src/simulation.jl:102-133`
* `generate_simulation_files(circ, prefix, n)`  -- `src/simulation.jl:52-77`
* `output_params_dict`                          -- `src/outputs.jl:47-78`

The contraction itself always runs on the GPU through libqxhip (`Executor`); there
is no CPU fallback.  Saved plans are `<prefix>.qx` (DSL text), `<prefix>.npz` (the
initial tensors; JLD2 is HDF5-based and not readable here) and `<prefix>.yml`.
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np

from . import _native
from .circuits import Circuit
from .compute_graph import build_compute_graph
from .planning import contraction_scheme, sweep_contraction_plan
from .tn import TensorNetworkCircuit, convert_to_tnc


def amplitudes_all(qubits: int):
    """All bitstrings of `qubits` bits, most significant first (`simulation.jl:14-16`)."""
    return (format(x, f"0{qubits}b") for x in range(2 ** qubits))


def amplitudes_uniform(qubits: int, seed: int | None, number_amplitudes: int):
    """`number_amplitudes` uniform random bitstrings with replacement
    (`simulation.jl:24-28`); numpy's PCG64 stream replaces Julia's MersenneTwister."""
    rng = np.random.Generator(np.random.PCG64(seed))
    draws = rng.integers(0, 2, size=(number_amplitudes, qubits))
    return ["".join("01"[b] for b in row) for row in draws]


_CTX: dict = {}


def context(device: int = 0) -> _native.Context:
    """Process-wide context per device (one HIP stream each)."""
    if device not in _CTX:
        _CTX[device] = _native.Context(device)
    return _CTX[device]


class Executor:
    """A saved plan (DSL + initial tensors) resident on one GPU."""

    def __init__(self, dsl: str, data: dict, precision: str = "c128", device: int = 0,
                 batch: int | None = None, arena_mb: int | None = None, graph: int | None = None,
                 profile: int | None = None, **options):
        """`options` are further `qx_plan_set_option` names (autotune, kernel,
        fuse_slices), set before finalize."""
        self.dsl = dsl
        self.precision = precision
        self.plan = _native.Plan(context(device), dsl, data, precision, batch=batch,
                                 arena_mb=arena_mb, graph=graph, profile=profile, **options)
        self.num_chars = self.plan.num_chars
        self.num_slices = self.plan.num_slices

    def amplitudes(self, bitstrings, slice_begin: int = 0, slice_end: int | None = None) -> np.ndarray:
        return self.plan.amplitudes(bitstrings, slice_begin, slice_end)

    def update_tensor(self, key: str, value) -> None:
        """New values of one data key after finalize (`qx_plan_update_tensor`)."""
        self.plan.update_tensor(key, value)

    def vjp(self, bitstrings, amp_cot) -> dict:
        return self.plan.vjp(bitstrings, amp_cot)

    def amplitudes_uniform(self, num: int, seed: int) -> tuple:
        return self.plan.amplitudes_uniform(num, seed)

    def sample_rejection(self, num_samples: int, seed: int, M: float = 1e-4, fix_M: bool = False,
                         max_candidates: int | None = None) -> tuple:
        return self.plan.sample_rejection(num_samples, seed, M, fix_M, max_candidates)

    def stats(self) -> dict:
        return self.plan.stats()

    def close(self):
        self.plan.close()


def single_amplitude(tnc: TensorNetworkCircuit, plan: list, amplitude: str | None = None,
                     precision: str = "c128", device: int = 0) -> complex:
    """Amplitude of one bitstring (`src/simulation.jl:86-91`).  The network's output
    tensors are bound to the bitstring's characters and the plan is contracted on the
    GPU; `tnc` itself is not modified."""
    n = tnc.num_qubits
    amplitude = amplitude or "0" * n
    if len(amplitude) != n:
        raise ValueError("bitstring length differs from qubit count")
    sim = tnc
    if not tnc.output_tensors:
        sim = tnc.copy()
        sim.add_output(amplitude)
    g = build_compute_graph(sim, plan)
    ex = Executor(g.to_dsl(), g.data, precision=precision, device=device, batch=1)
    try:
        return complex(ex.amplitudes([amplitude])[0])
    finally:
        ex.close()


def run_simulation(circ: Circuit, num_amplitudes: int | None = None, seed: int | None = None,
                   precision: str = "c128", device: int = 0) -> "OrderedDict[str, complex]":
    """Amplitudes of all bitstrings (n <= 30) or of `num_amplitudes` uniform ones
    (`src/simulation.jl:102-133`).  All unique bitstrings are contracted in batches
    in one call instead of one contraction per bitstring."""
    tnc = convert_to_tnc(circ)
    plan = sweep_contraction_plan(tnc)
    n = tnc.num_qubits
    if num_amplitudes is None and n > 30:
        num_amplitudes = 1000
    bits = list(amplitudes_all(n)) if num_amplitudes is None else amplitudes_uniform(n, seed, num_amplitudes)
    uniq = list(OrderedDict.fromkeys(bits))
    g = build_compute_graph(tnc, plan)
    ex = Executor(g.to_dsl(), g.data, precision=precision, device=device)
    try:
        amps = ex.amplitudes(uniq)
    finally:
        ex.close()
    return OrderedDict((b, complex(a)) for b, a in zip(uniq, amps))


def run_outputs(ex: Executor, output_args: dict):
    """Execute the output section of a saved plan's parameters (the `output` entry
    that `generate_parameter_file` writes, `src/simulation.jl:149-155`, built by
    `output_params_dict`, `src/outputs.jl:47-78`) on the GPU:

    * List      -> OrderedDict bitstring -> amplitude of the given bitstrings;
    * Uniform   -> OrderedDict of `num_samples` device-generated uniform bitstrings;
    * Rejection -> (list of sampled bitstrings, their amplitudes, final M, candidates).
    A missing seed is 0."""
    method = str(output_args["method"]).lstrip(":")
    p = output_args.get("params", {})
    seed = p.get("seed")
    seed = 0 if seed is None else int(seed)
    if method == "List":
        bits = list(p["bitstrings"])
        amps = ex.amplitudes(bits)
        return OrderedDict((b, complex(a)) for b, a in zip(bits, np.asarray(amps).reshape(len(bits), -1)[:, 0]))
    if method == "Uniform":
        bits, amps = ex.amplitudes_uniform(int(p["num_samples"]), seed)
        return OrderedDict((b, complex(a)) for b, a in zip(bits, np.asarray(amps).reshape(len(bits), -1)[:, 0]))
    if method == "Rejection":
        return ex.sample_rejection(int(p["num_samples"]), seed, float(p.get("M", 1e-4)), bool(p.get("fix_M", False)))
    raise ValueError(f'Output method "{method}" not supported')


def output_params_dict(num_qubits: int, num_outputs: int = 10, output_method: str = "List",
                       seed: int | None = None, M: float = 0.0001, fix_M: bool = False,
                       bitstrings: list | None = None) -> OrderedDict:
    """Output-sampling parameters (`src/outputs.jl:47-78`)."""
    out = OrderedDict(method=output_method)
    params = OrderedDict()
    if output_method == "Rejection":
        params.update(num_qubits=num_qubits, M=M, fix_M=fix_M, seed=seed, num_samples=num_outputs)
    elif output_method == "List":
        if bitstrings is None:
            bitstrings = amplitudes_uniform(num_qubits, seed, num_outputs)
        else:
            num_outputs = len(bitstrings)
        params.update(num_samples=num_outputs, bitstrings=list(bitstrings))
    elif output_method == "Uniform":
        params.update(num_qubits=num_qubits, num_samples=num_outputs, seed=seed)
    else:
        raise ValueError(f'Output method "{output_method}" not supported')
    out["params"] = params
    return out


def generate_simulation_files(circ: Circuit, output_prefix: str = "simulation_input",
                              number_bonds_to_slice: int = 2, decompose: bool = False,
                              seed: int | None = None, output_args=None, planner: str = "auto",
                              open_outputs=(), **kwargs) -> tuple:
    """Write `<prefix>.qx`, `<prefix>.npz`, `<prefix>.yml` (`src/simulation.jl:52-77`).

    planner: "sweep" (circuit-aware fused sweep, no slicing), "greedy" (randomised
    greedy + greedy slicing), or "auto" (sweep when nothing is sliced)."""
    import yaml

    tnc = convert_to_tnc(circ, decompose=decompose, open_outputs=open_outputs)
    if planner == "auto":
        planner = "sweep" if number_bonds_to_slice == 0 else "greedy"
    if planner == "sweep":
        plan = sweep_contraction_plan(tnc, **kwargs)
        groups, meta = [], OrderedDict([("Determination of contraction plan", OrderedDict(
            [("Method used", "fused gate sweep"), ("Hypergraph used", True)]))])
    else:
        groups, plan, meta = contraction_scheme(tnc, number_bonds_to_slice,
                                                seed=42 if seed is None else seed, **kwargs)
    g = build_compute_graph(tnc, plan, groups, metadata=meta)
    with open(output_prefix + ".qx", "w") as f:
        f.write(g.to_dsl())
    np.savez(output_prefix + ".npz", **{k: np.asarray(v) for k, v in g.data.items()})
    if output_args is None:
        output_args = output_params_dict(tnc.num_qubits, seed=seed)
    with open(output_prefix + ".yml", "w") as f:
        yaml.safe_dump({"output": _plain(output_args)}, f, sort_keys=False)
    return g, tnc, plan, groups


def _plain(x):
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    return x


def load_simulation_files(prefix: str) -> tuple:
    """(dsl text, {key: array}, params dict) of a saved plan."""
    import yaml

    with open(prefix + ".qx") as f:
        dsl = f.read()
    with np.load(prefix + ".npz", allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    params = {}
    try:
        with open(prefix + ".yml") as f:
            params = yaml.safe_load(f) or {}
    except FileNotFoundError:
        pass
    return dsl, data, params
