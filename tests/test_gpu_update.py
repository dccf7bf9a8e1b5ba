"""Tensor updates after finalize (qx_plan_update_tensor): the differentiable seam of a
variational loop.  New gate angles go into the same device buffers; amplitudes and
gradients must then equal a freshly built plan's oracle values at the new angles,
eagerly and when a captured graph is replayed.

Bars: complex128 amplitudes 1e-12 relative per amplitude against the oracle;
gradients 1e-11 normwise against the oracle's environment tensors and 1e-8 against
central finite differences (h = 1e-6) of the FP64 statevector; returning to the
first angles reproduces the first results bit for bit."""

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import Executor, circuits, gradients

pytestmark = pytest.mark.gpu

BITS = ["010110", "110001", "001111", "111000", "+-0101"]


def _rel(a, r):
    return float(np.max(np.abs(np.asarray(a) - r) / np.abs(r)))


@pytest.mark.parametrize("graph", [0, 1])
def test_param_update_amplitudes_and_gradients(graph):
    c = circuits.create_param_rqc_circuit(2, 3, 3, seed=11)
    pc = gradients.ParamCircuit(c, batch=8, graph=graph)
    th0 = c.params()
    a0 = pc.amplitudes(BITS)
    l0, g0, _ = pc.loss_and_grad(BITS)
    th1 = np.random.default_rng(3).uniform(0.0, 2 * np.pi, size=th0.size)
    pc.set_params(th1)
    c1 = c.with_params(th1)
    a1 = pc.amplitudes(BITS)
    # oracle on the plan's own DSL with the data the host now holds, and the statevector
    ref = O.amplitudes(pc.dsl, pc.graph.data, BITS)
    assert _rel(a1, ref) <= 1e-12
    assert _rel(a1, qxtest.statevector_amps(c1, BITS)) <= 1e-10
    assert _rel(a1, a0) > 1e-3                       # the update took effect
    # gradient of the loss at the new angles: per key against the environment tensors,
    # per angle against central differences of the statevector loss
    l1, g1, _ = pc.loss_and_grad(BITS)
    kg = pc.executor.vjp(BITS, 2.0 * a1)
    env = O.env_gradient(pc.dsl, pc.graph.data, BITS, 2.0 * ref)
    for k in env:
        assert np.linalg.norm(kg[k] - env[k]) <= 1e-11 * max(np.linalg.norm(env[k]), 1e-300), k
    h = 1e-6
    eye = np.eye(th1.size)
    pick = np.random.default_rng(4).choice(th1.size, 12, replace=False)
    for p in pick:
        lp = np.sum(np.abs(qxtest.statevector_amps(c.with_params(th1 + h * eye[p]), BITS)) ** 2)
        lm = np.sum(np.abs(qxtest.statevector_amps(c.with_params(th1 - h * eye[p]), BITS)) ** 2)
        assert abs(g1[p] - (lp - lm) / (2 * h)) < 1e-8 * max(1.0, abs(g1[p]))
    assert l1 == pytest.approx(float(np.sum(np.abs(ref) ** 2)), rel=1e-12)
    # back to the first angles: the first results again, bit for bit
    pc.set_params(th0)
    assert np.array_equal(pc.amplitudes(BITS), a0)
    l2, g2, _ = pc.loss_and_grad(BITS)
    assert l2 == l0 and np.array_equal(g2, g0)


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64_4m", 5e-5)])
def test_update_random_network_key(precision, tol):
    net = qxtest.RandomNet(7, n_tensors=8, n_indices=11, open_legs=0, dims=(2,))
    from qxhip import compute_graph, planning

    plan = planning.contraction_plan(net.tn, trials=2, seed=7)
    g = compute_graph.build_compute_graph(net, plan, dedupe=False)
    dsl = g.to_dsl()
    ex = Executor(dsl, g.data, precision=precision, graph=1)
    before = ex.amplitudes([""])
    rng = np.random.default_rng(8)
    data = dict(g.data)
    for key in list(data)[:3]:
        data[key] = rng.normal(size=data[key].shape) + 1j * rng.normal(size=data[key].shape)
        ex.update_tensor(key, data[key])
    after = ex.amplitudes([""])
    ref = O.amplitudes(dsl, data, [""])
    assert np.max(np.abs(after - ref)) <= tol * np.max(np.abs(ref))
    assert np.max(np.abs(after - before)) > 1e-3 * np.max(np.abs(ref))


def test_update_rejects_wrong_size_and_unknown_key():
    from qxhip._native import QxError

    c = circuits.create_rqc_circuit(2, 2, 3, seed=1)
    t, plan, g = qxtest.graph_for(c)
    ex = Executor(g.to_dsl(), g.data)
    key = next(iter(g.data))
    with pytest.raises(QxError) as e:
        ex.update_tensor(key, np.zeros(g.data[key].size + 1))
    assert e.value.status == -4
    with pytest.raises(QxError) as e:
        ex.update_tensor("no_such_key", np.zeros(4))
    assert e.value.status == -4
