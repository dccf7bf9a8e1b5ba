"""hipGraph replay of qx_amplitudes: every replay, with eager work and host
synchronisation between replays and new bitstrings each time, must equal the
eager result bit for bit (same kernels, same order)."""

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import Executor, circuits, simulation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("planner,nslice,nb,batch", [("sweep", 0, 16, 16), ("greedy", 2, 10, 4)])
def test_graph_replays_match_eager(planner, nslice, nb, batch):
    circ = circuits.create_rqc_circuit(3, 3, 10, seed=21)
    t, plan, g = qxtest.graph_for(circ, planner=planner, nslice=nslice)
    dsl = g.to_dsl()
    eager = Executor(dsl, g.data, batch=batch)
    graphed = Executor(dsl, g.data, batch=batch, graph=1)
    for rep in range(4):
        bits = simulation.amplitudes_uniform(9, 100 + rep, nb)
        want = eager.amplitudes(bits)
        got = graphed.amplitudes(bits)
        assert np.array_equal(got, want), f"replay {rep}"
        if rep == 0:
            ref = O.amplitudes(dsl, g.data, bits)
            assert np.max(qxtest.rel_err(got, ref)) < 1e-12
        _ = eager.amplitudes(bits[:1])          # eager work on the same device between replays
    part = graphed.amplitudes(bits, 0, 1)      # another slice range: another graph
    assert np.array_equal(part, eager.amplitudes(bits, 0, 1))


def test_call_beyond_graph_launch_cap_runs_uncaptured():
    """A call of more launches than the capture limit (256 looped slices x 8 batches
    of bitstrings, about 50K launches) runs without capture under the graph option,
    with the eager results bit for bit."""
    circ = circuits.create_rqc_circuit(3, 3, 10, seed=21)
    t, plan, g = qxtest.graph_for(circ, planner="greedy", nslice=8)
    dsl = g.to_dsl()
    bits = simulation.amplitudes_uniform(9, 7, 32)
    eager = Executor(dsl, g.data, batch=4, fuse_slices=0)
    graphed = Executor(dsl, g.data, batch=4, graph=1, fuse_slices=0)
    want = eager.amplitudes(bits)
    for rep in range(2):
        assert np.array_equal(graphed.amplitudes(bits), want), f"call {rep}"
    st = graphed.stats()
    assert st["last_call_launches"] > st["max_graph_launches"]
    fused = Executor(dsl, g.data, batch=4)
    assert np.max(qxtest.rel_err(want, fused.amplitudes(bits))) < 1e-12
    ref = O.amplitudes(dsl, g.data, bits[:2])
    assert np.max(qxtest.rel_err(want[:2], ref)) < 1e-12


@pytest.mark.parametrize("planner,nslice,nb,batch", [("sweep", 0, 8, 8), ("greedy", 2, 6, 4)])
def test_vjp_graph_replays_match_eager(planner, nslice, nb, batch):
    """qx_amplitudes_vjp replayed from its captured graph equals the eager pullback,
    with forward calls and other bitstrings between the replays."""
    circ = circuits.create_rqc_circuit(3, 3, 8, seed=5)
    t, plan, g = qxtest.graph_for(circ, planner=planner, nslice=nslice)
    dsl = g.to_dsl()
    eager = Executor(dsl, g.data, batch=batch)
    graphed = Executor(dsl, g.data, batch=batch, graph=1)
    for rep in range(3):
        bits = simulation.amplitudes_uniform(9, 200 + rep, nb)
        cot = np.random.default_rng(rep).normal(size=nb) + 1j
        want = eager.vjp(bits, cot)
        got = graphed.vjp(bits, cot)
        for k in want:
            assert np.array_equal(got[k], want[k]), (rep, k)
        _ = graphed.amplitudes(bits)             # forward graph between vjp replays


@pytest.mark.parametrize("name", ["rqc_3x3_d16", "rqc_5x5_d20"])
def test_small_level_walk_matches_level_launches(name):
    """The small steps' dependency levels walked by one workgroup in one launch
    (k_small_walk, plan option small_walk) give the one-launch-per-level results bit for
    bit (same chunks, same arithmetic), eagerly and on graph replay, and the pullback too."""
    from qxhip import workloads

    dsl, data, bits = workloads.WORKLOADS[name].load()
    bits = bits[:64]
    walk = Executor(dsl, data, batch=64, graph=1, small_walk=1)
    levels = Executor(dsl, data, batch=64)
    try:
        want = levels.amplitudes(bits)
        for _ in range(3):
            assert np.array_equal(walk.amplitudes(bits), want)
        kinds = {s["variant"] for s in next(b for b in walk.stats()["batches"] if b["active"])["steps"]}
        assert "k_small_walk" in kinds
        kinds = {s["variant"] for s in next(b for b in levels.stats()["batches"] if b["active"])["steps"]}
        assert "k_small_walk" not in kinds and "k_small_batch" in kinds
        if name == "rqc_3x3_d16":
            cot = np.linspace(0.5, 1.5, len(bits)) + 0.25j
            gw, gl = walk.vjp(bits, cot), levels.vjp(bits, cot)
            for k in gl:
                assert np.array_equal(gw[k], gl[k]), k
    finally:
        walk.close()
        levels.close()
