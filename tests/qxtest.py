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


def interleaved_gemm_program(nk=11, nm=8, nn=8, seed=0, view_a=False, a_data=False):
    """Two gate contractions that leave their results with the legs of one large
    contraction interleaved, then that contraction C = A.B over nk contracted legs:
    A carries contracted and A-only legs alternately, B contracted and B-only legs
    alternately (all dims 2).  view_a slices contracted label 1 through views of A and
    B (their layouts are then pinned); a_data makes A a data tensor holding its M legs
    then its K legs in a shuffled order (B must follow that order).
    Returns (dsl, data, reference C flattened column-major)."""
    rng = np.random.default_rng(seed)
    K = list(range(1, nk + 1))
    M = list(range(100, 100 + nm))
    N = list(range(200, 200 + nn))
    xa = [l for pair in zip(K, M) for l in pair] + K[len(M):] + M[len(K):]
    if a_data:   # A is the data tensor itself: its M run, then its K run in a shuffled order
        xa = M + list(rng.permutation(K))
    xb = [l for pair in zip(N, K) for l in pair] + N[len(K):] + K[len(N):]
    # A = X . G on A's first contracted leg (label 1 takes the gate's input label 900)
    xa_in = [900 if l == 1 else l for l in xa]
    xb_in = [901 if l == 2 else l for l in xb]
    X = rng.normal(size=[2] * len(xa)) + 1j * rng.normal(size=[2] * len(xa))
    Y = rng.normal(size=[2] * len(xb)) + 1j * rng.normal(size=[2] * len(xb))
    G1 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    G2 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    cl = M + N
    fmt = lambda ls: ",".join(map(str, ls))
    lines = ["# version: 0.4.0",
             f"load X data_x {fmt([2] * len(xa))}", f"load Y data_y {fmt([2] * len(xb))}",
             "load G1 data_g1 2,2", "load G2 data_g2 2,2",
             f"ncon A {fmt(xa)} X {fmt(xa_in)} G1 900,1",
             f"ncon B {fmt(xb)} Y {fmt(xb_in)} G2 901,2"]
    if a_data:
        xa = [int(l) for l in xa]
        X = rng.normal(size=[2] * len(xa)) + 1j * rng.normal(size=[2] * len(xa))
        lines = [lines[0], f"load A data_x {fmt([2] * len(xa))}", lines[2], lines[4], lines[6]]
    if view_a:   # slice contracted label 1 in both operands
        lines.append(f"view Av A v1 {xa.index(1) + 1} 2")
        lines.append(f"view Bv B v1 {xb.index(1) + 1} 2")
    lines += [f"ncon C {fmt(cl)} {'Av' if view_a else 'A'} {fmt(xa)} {'Bv' if view_a else 'B'} {fmt(xb)}",
              "save output C"]
    data = {"data_x": X, "data_y": Y, "data_g1": G1, "data_g2": G2}
    if a_data:
        del data["data_g1"]
    A = X if a_data else O.ncon(X, xa_in, G1, [900, 1], xa)
    B = O.ncon(Y, xb_in, G2, [901, 2], xb)
    ref = O.ncon(A, xa, B, xb, cl).reshape(-1, order="F")
    return "\n".join(lines) + "\n", data, ref


def users_guide_ghz2():
    """(dsl, data) of the reference's DSL example (`docs/src/users_guide.md:71-90`,
    kept as `tests/golden/users_guide_ghz2.qx`): two-qubit GHZ (H on qubit 1, CNOT
    with control 1 and target 2) with both slice variables.  Its data keys are built
    here from the gates, following the program's leg use:
    * data_1 (2): the input state |0> (t4 and t5 are the two inputs);
    * data_2 (2,2): H, leg 1 its output (the control wire's hyperindex), leg 2 its input;
    * data_3 (2,2,2): the CNOT's target part, [out, in, c] = (X^c)[out, in], leg 3 on
      the control's hyperindex (the CNOT is diagonal on its control);
    * data_4 (2): the control part on the hyperindex, all ones."""
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "users_guide_ghz2.qx")
    with open(path) as f:
        dsl = f.read()
    h = circuits.gate_matrix(circuits.Gate("h", (1,)))
    x = circuits.gate_matrix(circuits.Gate("x", (1,)))
    cnot_t = np.stack([np.eye(2, dtype=np.complex128), x], axis=2)
    data = {"data_1": np.array([1.0, 0.0], dtype=np.complex128), "data_2": h, "data_3": cnot_t,
            "data_4": np.ones(2, dtype=np.complex128)}
    return dsl, data
