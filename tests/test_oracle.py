"""Pins of the CPU FP64 oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle (dropped term,
wrong leg order, transposed operand, wrong slice arithmetic) fails it:
* reference known-answer vectors: GHZ-3 statevector (test/test_tn_conversion.jl:38-40,
  test/test_contraction_planning.jl:60) and GHZ-5 amplitudes (test/test_simulation.jl:24-25);
* an independent FP64 statevector simulator, itself pinned by closed forms;
* np.einsum on random steps of every label class;
* Sum |amp|^2 = 1 and the slice-sum identity.
"""

import itertools
import math

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import circuits, simulation

S2 = 1 / math.sqrt(2)


# --------------------------------------------------------------------- statevector pins
def test_statevector_closed_forms():
    n = 4
    # X on qubits 1 and 3: basis state 1010
    psi = O.statevector(n, [(circuits.X, [1]), (circuits.X, [3])])
    assert O.sv_amplitude(psi, "1010") == 1.0
    assert O.sv_amplitude(psi, "1000") == 0.0
    # Hadamard layer: uniform 2^{-n/2}
    psi = O.statevector(n, [(circuits.H, [q]) for q in range(1, n + 1)])
    for b in simulation.amplitudes_all(n):
        assert O.sv_amplitude(psi, b) == pytest.approx(0.25, abs=1e-15)
    # CX with [target, ctrl] order: H(2) CX(target 1, ctrl 2) -> (|00> + |11>)/sqrt2
    psi = O.statevector(2, [(circuits.H, [2]), (circuits.CX, [1, 2])])
    assert O.sv_amplitude(psi, "11") == pytest.approx(S2)
    assert O.sv_amplitude(psi, "01") == 0.0
    # a CX read with the opposite qubit order would give |01>+|11> instead
    psi_wrong = O.statevector(2, [(circuits.H, [2]), (circuits.CX, [2, 1])])
    assert O.sv_amplitude(psi_wrong, "11") == 0.0


def test_statevector_plus_minus_and_norm():
    circ = circuits.create_rqc_circuit(2, 2, 5, seed=1)
    psi = O.statevector(4, qxtest.circuit_gates(circ))
    assert np.sum(np.abs(psi) ** 2) == pytest.approx(1.0, abs=1e-14)
    a0, a1 = O.sv_amplitude(psi, "0010"), O.sv_amplitude(psi, "1010")
    assert O.sv_amplitude(psi, "+010") == pytest.approx(S2 * (a0 + a1))
    assert O.sv_amplitude(psi, "-010") == pytest.approx(S2 * (a0 - a1))


def test_gate_matrices():
    for g in (circuits.SQRT_X, circuits.SQRT_Y, circuits.SQRT_W):
        assert np.allclose(g @ g.conj().T, np.eye(2))
    assert np.allclose(circuits.SQRT_X @ circuits.SQRT_X, circuits.X)
    assert np.allclose(circuits.SQRT_Y @ circuits.SQRT_Y, circuits.Y)
    assert np.allclose(circuits.SQRT_W @ circuits.SQRT_W, circuits.W)
    th = 0.37
    h = 1e-6
    assert np.allclose((circuits.rz(th + h) - circuits.rz(th - h)) / (2 * h), circuits.drz(th), atol=1e-9)
    assert np.allclose((circuits.ry(th + h) - circuits.ry(th - h)) / (2 * h), circuits.dry(th), atol=1e-9)


# --------------------------------------------------------------------- reference KATs
@pytest.mark.parametrize("decompose", [False, True])
@pytest.mark.parametrize("planner", ["greedy", "sweep"])
def test_ghz3_kat(decompose, planner):
    circ = circuits.create_test_circuit()
    t, plan, g = qxtest.graph_for(circ, planner=planner, decompose=decompose, no_output=True)
    out = O.amplitudes(g.to_dsl(), g.data, [""])[0].reshape(-1, order="F")
    ref = np.zeros(8)
    ref[[0, 7]] = S2
    assert np.allclose(out, ref, atol=1e-15)


def test_ghz5_kat():
    circ = circuits.create_ghz_circuit(5)
    t, plan, g = qxtest.graph_for(circ)
    bits = list(simulation.amplitudes_all(5))
    amps = O.amplitudes(g.to_dsl(), g.data, bits)
    assert amps[0] == pytest.approx(S2) and amps[-1] == pytest.approx(S2)
    assert np.allclose(amps[1:-1], 0)


# --------------------------------------------------------------------- executor vs statevector
@pytest.mark.parametrize("planner", ["sweep", "greedy"])
def test_oracle_matches_statevector_rqc(planner):
    circ = circuits.create_rqc_circuit(3, 3, 8, seed=42)
    t, plan, g = qxtest.graph_for(circ, planner=planner)
    bits = simulation.amplitudes_uniform(9, 4242, 24) + ["+-+-+-+-0"]
    amps = O.amplitudes(g.to_dsl(), g.data, bits)
    sv = qxtest.statevector_amps(circ, bits)
    assert np.max(np.abs(amps - sv)) < 1e-14


def test_norm_over_all_bitstrings():
    circ = circuits.create_rqc_circuit(2, 3, 6, seed=5)
    t, plan, g = qxtest.graph_for(circ)
    amps = O.amplitudes(g.to_dsl(), g.data, list(simulation.amplitudes_all(6)))
    assert np.sum(np.abs(amps) ** 2) == pytest.approx(1.0, abs=1e-13)


def test_planted_transpose_fails_statevector_pin():
    """Transposing one gate's tensor in the saved data must break agreement with the
    statevector: the pin sees leg-order mistakes."""
    circ = circuits.create_rqc_circuit(2, 2, 4, seed=2)
    t, plan, g = qxtest.graph_for(circ)
    bits = list(simulation.amplitudes_all(4))
    sv = qxtest.statevector_amps(circ, bits)
    data = dict(g.data)
    key = next(k for k, v in data.items() if v.shape == (2, 2) and not np.allclose(v, v.T))
    data[key] = data[key].T
    amps = O.amplitudes(g.to_dsl(), data, bits)
    assert np.max(np.abs(amps - sv)) > 1e-3


# --------------------------------------------------------------------- ncon vs einsum
def _einsum_ref(A, al, B, bl, cl):
    letters = {l: chr(97 + i) for i, l in enumerate(sorted(set(al + bl)))}
    spec = "".join(letters[l] for l in al) + "," + "".join(letters[l] for l in bl) + "->" + \
        "".join(letters[l] for l in cl)
    return np.einsum(spec, A, B)


@pytest.mark.parametrize("seed", range(40))
def test_ncon_vs_einsum(seed):
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(1, 8))
    labs = list(range(1, nl + 1))
    dims = {l: int(rng.choice([1, 2, 3])) for l in labs}
    cls = {l: rng.choice(["H", "K", "M", "N"]) for l in labs}
    al = [l for l in labs if cls[l] in "HKM"]
    bl = [l for l in labs if cls[l] in "HKN"]
    rng.shuffle(al)
    rng.shuffle(bl)
    cl = [l for l in labs if cls[l] in "HMN"]
    rng.shuffle(cl)
    A = rng.normal(size=[dims[l] for l in al]) + 1j * rng.normal(size=[dims[l] for l in al])
    B = rng.normal(size=[dims[l] for l in bl]) + 1j * rng.normal(size=[dims[l] for l in bl])
    C = O.ncon(np.asarray(A), al, np.asarray(B), bl, cl)
    R = _einsum_ref(np.asarray(A), al, np.asarray(B), bl, cl)
    assert C.shape == R.shape
    assert np.allclose(C, R, rtol=1e-13, atol=1e-13)


def test_ncon_rejects_invalid_labels():
    A = np.ones((2, 2))
    with pytest.raises(ValueError):
        O.ncon(A, [1, 1], A, [1, 2], [2])          # repeated label
    with pytest.raises(ValueError):
        O.ncon(A, [1, 2], A, [2, 3], [3])          # label 1 summed out of A alone


# --------------------------------------------------------------------- slicing and summation
def test_slice_sum_identity_and_enumeration():
    circ = circuits.create_rqc_circuit(3, 3, 8, seed=9)
    t, plan, g = qxtest.graph_for(circ, planner="greedy", nslice=3)
    t0, plan0, g0 = qxtest.graph_for(circ, planner="greedy", nslice=0)
    dsl = g.to_dsl()
    cmds = O.parse(dsl)
    vs = O.slice_vars(cmds)
    assert [v for v, _ in vs] == ["v1", "v2", "v3"]
    assigns = list(O.slice_assignments(cmds))
    assert len(assigns) == 8
    assert [a["v1"] for a in assigns[:4]] == [0, 1, 0, 1]     # first variable fastest
    assert [a["v3"] for a in assigns] == [0] * 4 + [1] * 4
    bits = simulation.amplitudes_uniform(9, 3, 6)
    sliced = O.amplitudes(dsl, g.data, bits)
    full = O.amplitudes(g0.to_dsl(), g0.data, bits)
    assert np.allclose(sliced, full, rtol=1e-13, atol=1e-16)
    per = O.per_slice(dsl, g.data, bits[0])
    assert abs(sum(per) - sliced[0]) < 1e-15


def test_neumaier_recovers_cancellation():
    acc = O.Neumaier()
    for x in [1.0, 1e100, 1.0, -1e100]:
        acc.add(x)
    assert acc.value() == 2.0
    acc = O.Neumaier((2,))
    for x in itertools.repeat(np.array([0.1 + 0.1j, 1e-17]), 10):
        acc.add(x)
    assert acc.value()[0] == pytest.approx(1 + 1j, abs=1e-16)


# --------------------------------------------------------------------- gradient oracle
def test_env_gradient_matches_finite_differences():
    """Environment-tensor gradients chained to gate angles equal central finite
    differences of the statevector loss L = sum |amp|^2 (h = 1e-6, FP64)."""
    from qxhip import compute_graph, planning, tn

    c = circuits.create_param_rqc_circuit(2, 2, 2, seed=1)
    tnc, names = tn.gate_tensor_map(c)
    tnc.add_input()
    tnc.add_output()
    g = compute_graph.build_compute_graph(tnc, planning.sweep_contraction_plan(tnc), dedupe=False)
    bits = ["0101", "1100", "+0-1"]
    dsl = g.to_dsl()
    amps = O.amplitudes(dsl, g.data, bits)
    kg = O.env_gradient(dsl, g.data, bits, 2 * amps)
    key_of = {cm.name: cm.args[0] for cm in g.commands if cm.kind == "load"}
    th = c.params()
    grad = np.zeros(len(th))
    for gg, nm in zip(c.ops, names):
        if gg.param is None:
            continue
        s = tn.TensorNetworkCircuit(4)
        dn = s.push(circuits.gate_qubits(gg), circuits.gate_matrix_derivative(gg))
        d = s.tn[dn[0]].data
        grad[gg.param] += np.real(np.sum(np.conj(kg[key_of[nm[0]]].reshape(d.shape, order="F")) * d))

    def loss(t):
        return float(np.sum(np.abs(qxtest.statevector_amps(c.with_params(t), bits)) ** 2))

    h = 1e-6
    eye = np.eye(len(th))
    fd = np.array([(loss(th + h * eye[p]) - loss(th - h * eye[p])) / (2 * h) for p in range(len(th))])
    assert np.max(np.abs(grad - fd)) < 1e-8


def test_users_guide_dsl_example():
    """The reference's own DSL example (two-qubit GHZ, `docs/src/users_guide.md:71-90`)
    with data keys built from its gates: every bitstring over {0,1,+,-}^2 equals the
    FP64 statevector of H(1), CNOT(1->2); "00" and "11" are 1/sqrt2, "01" and "10" 0."""
    import itertools

    dsl, data = qxtest.users_guide_ghz2()
    bits = ["".join(p) for p in itertools.product("01+-", repeat=2)]
    out = O.amplitudes(dsl, data, bits)
    ref = qxtest.statevector_amps(circuits.create_ghz_circuit(2), bits)
    assert np.max(np.abs(out - ref)) < 1e-15
    s2 = 1 / np.sqrt(2)
    assert np.allclose(out[[0, 1, 4, 5]], [s2, 0, 0, s2], rtol=0, atol=1e-16)   # 00 01 10 11
    assert len(O.slice_vars(O.parse(dsl))) == 2


def test_step_shapes_hand_counted():
    """The oracle's per-step (B, M, N, K) on a program whose classes are counted by hand:
    label 2 is in A, B and C (batch, 4), label 1 only in A and C (M, 2), label 3 only in
    B and C (N, 16), label 4 in A and B only (K, 8); the second step sums labels 3 and 2
    (K = 64) and keeps label 1 (M = 2).  Exchanging M and N, or K and B, fails the
    tuple; the library's lowering (an independent classification in C++) agrees on B
    and K and on the pair {M, N} (it may exchange the operands)."""
    from qxhip import _native
    dsl = """# version: 0.4.0
load a data_1 2,4,8
load b data_2 8,4,16
ncon c 1,3,2 a 1,2,4 b 4,2,3
load f data_3 16,4
ncon e 1 c 1,3,2 f 3,2
save output e
"""
    shapes = O.step_shapes(dsl)
    assert shapes == [(4, 2, 16, 8), (1, 2, 1, 64)]
    # multiply-adds by brute force over every label assignment of the first step
    dims = {1: 2, 2: 4, 3: 16, 4: 8}
    macs = sum(1 for _ in itertools.product(*(range(dims[l]) for l in (1, 2, 3, 4))))
    B, M, N, K = shapes[0]
    assert macs == B * M * N * K
    info = _native.inspect(dsl, batch=1)
    lowered = [(s["B"], s["M"], s["N"], s["K"]) for s in info["steps"]]
    for (b, m, n, k), (lb, lm, ln, lk) in zip(shapes, lowered):
        assert (lb, lk) == (b, k) and sorted((lm, ln)) == sorted((m, n))
