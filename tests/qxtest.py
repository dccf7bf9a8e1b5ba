"""Shared helpers of the test suite: workload builders and error metrics."""

from __future__ import annotations

import numpy as np

import qxoracle as O
from qxhip import circuits, compute_graph, planning, tn


def circuit_gates(circ):
    return [(circuits.gate_matrix(g), circuits.gate_qubits(g)) for g in circ.ops]


def statevector_amps(circ, bitstrings):
    psi = O.statevector(circ.num_qubits, circuit_gates(circ))
    return O.sv_amplitudes(psi, bitstrings)


def graph_for(circ, planner="sweep", nslice=0, decompose=False, no_output=False, seed=42, **kw):
    t = tn.convert_to_tnc(circ, decompose=decompose, no_output=no_output)
    if planner == "sweep":
        plan = planning.sweep_contraction_plan(t, **kw)
        groups = []
    else:
        plan = planning.contraction_plan(t, trials=kw.get("trials", 8), seed=seed)
        groups = planning.expand_slice_bonds_to_hyperindices(
            t, planning.greedy_slicing(t, plan, nslice)) if nslice else []
    g = compute_graph.build_compute_graph(t, plan, groups)
    return t, plan, g


class RandomNet:
    """A random closed (or partly open) network for executor tests."""

    def __init__(self, seed, n_tensors=8, n_indices=12, max_rank=5, hyper=True, open_legs=0, dims=(2,)):
        rng = np.random.default_rng(seed)
        self.tn = tn.TensorNetwork()
        idx = [self.tn.new_index(int(rng.choice(dims))) for _ in range(n_indices)]
        legs = [[] for _ in range(n_tensors)]
        for i in idx:
            k = 3 if (hyper and rng.uniform() < 0.25) else 2
            for t in rng.choice(n_tensors, size=min(k, n_tensors), replace=False):
                if len(legs[t]) < max_rank:
                    legs[t].append(i)
        self.open = []
        for _ in range(open_legs):
            o = self.tn.new_index(2)
            legs[int(rng.integers(n_tensors))].append(o)
            self.open.append(o)
        # every index must have >= 2 holders unless open; drop singletons
        holders = {}
        for t, l in enumerate(legs):
            for i in l:
                holders.setdefault(i, []).append(t)
        for t in range(n_tensors):
            legs[t] = [i for i in legs[t] if len(holders[i]) >= 2 or i in self.open]
            shape = tuple(self.tn.dims[i] for i in legs[t])
            data = rng.normal(size=shape) + 1j * rng.normal(size=shape)
            self.tn.push(legs[t], data)
        self.num_qubits = 0
        self.output_tensors = []


def rel_err(a, r):
    a = np.asarray(a)
    r = np.asarray(r)
    return np.abs(a - r) / np.maximum(np.abs(r), 1e-300)


def gate_sequence_program(n: int, gates: list, seed: int = 0):
    """DSL of dense gates applied to a random 2^n tensor: gate j (k legs, a
    2^k x 2^k block) replaces the tensor's legs `gates[j]` (qubit positions) with
    new ones; the result keeps all n legs.  Returns (dsl, data)."""
    rng = np.random.default_rng(seed)
    qlab = list(range(1, n + 1))
    order = list(range(1, n + 1))
    nxt = n + 1
    lines = ["# version: 0.4.0", "load X data_x " + ",".join(["2"] * n)]
    data = {"data_x": rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)}
    cur = "X"
    for j, legs in enumerate(gates):
        k = len(legs)
        data[f"data_g{j}"] = (rng.normal(size=4 ** k) + 1j * rng.normal(size=4 ** k)) / 2 ** k
        lines.append(f"load G{j} data_g{j} " + ",".join(["2"] * (2 * k)))
        ins = [qlab[q] for q in legs]
        outs = list(range(nxt, nxt + k))
        nxt += k
        new = [l for l in order if l not in ins] + outs
        lines.append(f"ncon I{j} {','.join(map(str, new))} {cur} {','.join(map(str, order))} "
                     f"G{j} {','.join(map(str, ins + outs))}")
        for q, o in zip(legs, outs):
            qlab[q] = o
        order = new
        cur = f"I{j}"
    lines.append(f"save output {cur}")
    return "\n".join(lines) + "\n", data
