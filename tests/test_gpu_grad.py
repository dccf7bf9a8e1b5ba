"""GPU gradients (qx_amplitudes_vjp): against the oracle's environment-tensor
gradients (exact, per data key) and against central finite differences on gate
angles.  Tolerances: complex128 1e-11 normwise; complex64 1e-4."""

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import Executor, circuits, compute_graph, gradients, planning, tn

pytestmark = pytest.mark.gpu


def _nerr(a, r):
    return np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300)


def _check(ex, dsl, data, bits, cot, tol):
    g = ex.vjp(bits, cot)
    ref = O.env_gradient(dsl, data, bits, cot)
    for k in ref:
        assert _nerr(g[k], ref[k]) < tol, k


@pytest.mark.parametrize("seed", range(4))
def test_vjp_random_networks(seed):
    net = qxtest.RandomNet(seed, n_tensors=6 + seed, n_indices=8 + seed, open_legs=seed % 2,
                           dims=(2,) if seed < 2 else (2, 3))
    plan = planning.contraction_plan(net.tn, trials=2, seed=seed, open_indices=net.open)
    g = compute_graph.build_compute_graph(net, plan)
    dsl = g.to_dsl()
    ex = Executor(dsl, g.data)
    rng = np.random.default_rng(seed)
    R = ex.plan.result_elems
    cot = rng.normal(size=R) + 1j * rng.normal(size=R)
    _check(ex, dsl, g.data, [""], cot, 1e-11)


@pytest.mark.parametrize("nb,batch", [(1, 1), (3, 4), (5, 2)])
def test_vjp_rqc_batches(nb, batch):
    circ = circuits.create_rqc_circuit(2, 2, 4, seed=9)
    t, plan, g = qxtest.graph_for(circ)
    dsl = g.to_dsl()
    bits = ["0101", "1110", "+-01", "0000", "1011"][:nb]
    ex = Executor(dsl, g.data, batch=batch)
    rng = np.random.default_rng(nb)
    cot = rng.normal(size=nb) + 1j * rng.normal(size=nb)
    _check(ex, dsl, g.data, bits, cot, 1e-11)


def test_vjp_sliced_plan():
    circ = circuits.create_rqc_circuit(2, 2, 5, seed=4)
    t, plan, g = qxtest.graph_for(circ, planner="greedy", nslice=2)
    dsl = g.to_dsl()
    assert "view" in dsl
    bits = ["0110", "1001"]
    ex = Executor(dsl, g.data, batch=2)
    cot = np.array([0.3 - 0.2j, -1.1 + 0.5j])
    _check(ex, dsl, g.data, bits, cot, 1e-11)


def test_param_gradient_vs_finite_differences():
    c = circuits.create_param_rqc_circuit(2, 2, 3, seed=5)
    pc = gradients.ParamCircuit(c, batch=4)
    bits = ["0101", "1100", "0011", "1111"]
    loss, grad, amps = pc.loss_and_grad(bits)
    th = c.params()
    h = 1e-6
    eye = np.eye(len(th))
    fd = []
    for p in range(len(th)):
        lp = float(np.sum(np.abs(qxtest.statevector_amps(c.with_params(th + h * eye[p]), bits)) ** 2))
        lm = float(np.sum(np.abs(qxtest.statevector_amps(c.with_params(th - h * eye[p]), bits)) ** 2))
        fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)
    assert np.max(np.abs(grad - fd)) < 1e-8 * max(1.0, np.max(np.abs(fd)))
    assert loss == pytest.approx(float(np.sum(np.abs(qxtest.statevector_amps(c, bits)) ** 2)), rel=1e-12)


def test_param_gradient_c64():
    c = circuits.create_param_rqc_circuit(2, 2, 3, seed=6)
    bits = ["0101", "1100"]
    g64 = gradients.ParamCircuit(c, precision="c128").loss_and_grad(bits)[1]
    g32 = gradients.ParamCircuit(c, precision="c64_4m").loss_and_grad(bits)[1]
    assert _nerr(g32, g64) < 1e-4


@pytest.mark.parametrize("planner,nslice,nb,batch,prec", [("sweep", 0, 5, 2, "c128"), ("greedy", 2, 6, 4, "c128"),
                                                          ("greedy", 2, 4, 4, "c64_4m")])
def test_norm_vjp_equals_amplitudes_then_vjp(planner, nslice, nb, batch, prec):
    """qx_norm_vjp (the loss sum |amp|^2 in one pass, cotangent 2 amp formed on the
    device) against qx_amplitudes followed by qx_amplitudes_vjp of 2 amp, eager and
    replayed from its graph, with other calls between the replays; the gradient also
    against the oracle's environment tensors."""
    circ = circuits.create_rqc_circuit(2, 3, 5, seed=12)
    t, plan, g = qxtest.graph_for(circ, planner=planner, nslice=nslice)
    dsl = g.to_dsl()
    bits = circuits_bits = ["010110", "111000", "+-0101", "000000", "101011", "0110-+"][:nb]
    tol = 1e-11 if prec == "c128" else 1e-4
    for graph in (0, 1):
        ex = Executor(dsl, g.data, batch=batch, precision=prec, graph=graph)
        amps_ref = ex.amplitudes(bits)
        grad_ref = ex.vjp(bits, 2.0 * amps_ref)
        for rep in range(2):
            amps, grad = ex.plan.norm_vjp(bits)
            assert _nerr(amps, amps_ref) < (1e-12 if prec == "c128" else 1e-5), (graph, rep)
            for k in grad_ref:
                assert _nerr(grad[k], grad_ref[k]) < tol, (graph, rep, k)
            _ = ex.amplitudes(circuits_bits[:1])
        if prec == "c128" and graph == 0:
            ref = O.env_gradient(dsl, g.data, bits, 2.0 * O.amplitudes(dsl, g.data, bits))
            for k in ref:
                assert _nerr(grad[k], ref[k]) < 1e-11, k
        ex.close()


def test_norm_vjp_refuses_looped_slices():
    """With slice variables looped (fuse_slices 0) the full amplitudes are known only
    after every slice: the norm pullback reports QX_ERR_UNSUPPORTED and the parameter
    gradient falls back to amplitudes followed by the pullback, with the same result."""
    from qxhip import _native
    circ = circuits.create_rqc_circuit(2, 2, 5, seed=4)
    t, plan, g = qxtest.graph_for(circ, planner="greedy", nslice=2)
    dsl = g.to_dsl()
    bits = ["0110", "1001"]
    ex = Executor(dsl, g.data, batch=2, fuse_slices=0)
    with pytest.raises(_native.QxError) as e:
        ex.plan.norm_vjp(bits)
    assert e.value.status == _native.QX_ERR_UNSUPPORTED
    amps = ex.amplitudes(bits)
    ref = O.env_gradient(dsl, g.data, bits, 2.0 * amps)
    grad = ex.vjp(bits, 2.0 * amps)
    for k in ref:
        assert _nerr(grad[k], ref[k]) < 1e-11, k
    ex.close()


@pytest.mark.parametrize("planner,nslice", [("sweep", 0), ("greedy", 2)])
def test_param_gradient_one_pass_equals_two_calls(planner, nslice):
    """loss_and_grad's one-pass path (qx_norm_vjp, flat gradient gathered by one
    index) against the two calls (amplitudes, then the pullback of 2 amp with the
    per-key chain rule)."""
    c = circuits.create_param_rqc_circuit(2, 3, 3, seed=8)
    pc = gradients.ParamCircuit(c, planner=planner, nslice=nslice, batch=4)
    bits = ["010110", "110001", "001111", "111000", "+-0101"]
    loss, grad, amps = pc.loss_and_grad(bits)
    amps2 = pc.amplitudes(bits)
    grad2 = pc.param_grad(pc.executor.vjp(bits, 2.0 * amps2))
    assert _nerr(amps, amps2) < 1e-12
    assert loss == pytest.approx(float(np.sum(np.abs(amps2) ** 2)), rel=1e-12)
    assert _nerr(grad, grad2) < 1e-11
