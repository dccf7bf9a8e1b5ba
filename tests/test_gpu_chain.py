"""Chain passes (runs of gate-block steps applied to one large tensor in one pass
through LDS) against the oracle, the statevector and the unchained lowering."""

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import Executor, _native, circuits, simulation

pytestmark = pytest.mark.gpu

CIRCUITS = [(4, 4, 8), (3, 5, 10), (4, 4, 12), (2, 7, 10)]


@pytest.mark.parametrize("shape", CIRCUITS)
def test_chain_plan_matches_oracle_and_statevector(shape):
    r, c, d = shape
    circ = circuits.create_rqc_circuit(r, c, d, seed=7)
    t, plan, g = qxtest.graph_for(circ)
    dsl = g.to_dsl()
    info = _native.inspect(dsl, batch=16)
    assert any(s["kernel"] == "chain" for s in info["steps"])
    n = r * c
    bits = simulation.amplitudes_uniform(n, 31, 40)
    ex = Executor(dsl, g.data, batch=16)
    out = ex.amplitudes(bits)
    act = [b for b in ex.stats()["batches"] if b["active"]][0]
    assert any(s["kernel"] == "chain" for s in act["steps"])
    ref = O.amplitudes(dsl, g.data, bits)
    assert np.max(qxtest.rel_err(out, ref)) <= 1e-12
    sv = qxtest.statevector_amps(circ, bits)
    assert np.max(qxtest.rel_err(out, sv)) <= 1e-10
    plain = Executor(dsl, g.data, batch=16, chains=0)
    assert not any(s["kernel"] == "chain" for s in [b for b in plain.stats()["batches"]][0]["steps"])
    assert np.max(qxtest.rel_err(plain.amplitudes(bits), out)) <= 1e-12


def test_chain_plan_complex64():
    circ = circuits.create_rqc_circuit(4, 4, 10, seed=9)
    t, plan, g = qxtest.graph_for(circ)
    dsl = g.to_dsl()
    bits = simulation.amplitudes_uniform(16, 3, 32)
    a64 = Executor(dsl, g.data, precision="c64_4m", batch=32).amplitudes(bits)
    ref = O.amplitudes(dsl, g.data, bits)
    assert np.linalg.norm(a64 - ref) / np.linalg.norm(ref) < 5e-6


def test_chain_with_graph_replay_and_autotune():
    circ = circuits.create_rqc_circuit(3, 5, 10, seed=7)
    t, plan, g = qxtest.graph_for(circ)
    dsl = g.to_dsl()
    bits = simulation.amplitudes_uniform(15, 8, 24)
    ref = O.amplitudes(dsl, g.data, bits)
    ex = Executor(dsl, g.data, batch=8, graph=1, autotune=1)
    for _ in range(3):
        assert np.max(qxtest.rel_err(ex.amplitudes(bits), ref)) <= 1e-12
