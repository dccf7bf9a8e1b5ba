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


GATE_RUNS = {
    "k1-k5 mixed": [(0,), (1, 2), (3, 4, 5), (6, 7, 8, 9), (2, 3, 4, 5, 6)],
    "k4 x3 overlapping": [(10, 11, 12, 13), (13, 14, 15, 0), (0, 1, 2, 3)],
    "k5 x2 high legs": [(11, 12, 13, 14, 15), (6, 7, 8, 9, 10)],
    "k2 x4 low legs": [(0, 1), (1, 2), (2, 3), (0, 3)],
    # merged gate blocks: two 2-leg blocks into one 4-leg step, a 1-leg block absorbed
    # into a 4-leg one, a block moved back past a block on other legs, three blocks
    "merge pair and absorb": [(0, 1), (2, 3), (6, 7, 8, 9), (8,)],
    "merge past other legs": [(0, 5), (6, 7, 8, 9), (3, 4), (0, 3)],
    "merge three": [(0, 3), (5, 0), (6, 3), (9, 10, 11, 12), (12,), (1, 2)],
}

# chain steps after merging (the kernel's passes over the tile) for the runs above
MERGED_STEPS = {"merge pair and absorb": 2, "merge past other legs": 2, "merge three": 3, "k2 x4 low legs": 1}


@pytest.mark.parametrize("name", sorted(GATE_RUNS))
@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64_4m", 2e-5)])
def test_chain_gate_runs(name, precision, tol):
    """Every contracted-leg count (vector path for 1-8 values, matrix cores for 16
    and 32) on a 2^16 tensor, against the oracle and the unchained lowering."""
    dsl, data = qxtest.gate_sequence_program(16, GATE_RUNS[name], seed=len(name))
    info = _native.inspect(dsl, precision=precision)
    assert [s["kernel"] for s in info["steps"]] == ["chain"] * len(info["steps"])
    ex = Executor(dsl, data, precision=precision, batch=1)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    ref = O.amplitudes(dsl, data, [""])[0].reshape(-1, order="F")
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol
    plain = Executor(dsl, data, precision=precision, batch=1, chains=0)
    outp = np.asarray(plain.amplitudes([""])).reshape(-1)
    assert np.linalg.norm(out - outp) / np.linalg.norm(outp) < tol
    if name in MERGED_STEPS:
        ch = [s for s in info["steps"] if s["kernel"] == "chain"]
        assert sum(len(s["chain_tile"]["q"]) for s in ch) == MERGED_STEPS[name], [s["chain_tile"]["q"] for s in ch]
        assert sum(s["fused_steps"] for s in ch) == len(GATE_RUNS[name])


@pytest.mark.parametrize("variant", [0, 1, 2, 4, 5, 6, 8, 9, 10, 14])
@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64_4m", 2e-5)])
def test_chain_variants(variant, precision, tol):
    """Every path of the 16-value steps (vector FMAs, matrix cores with four or three
    real products per complex product) with plain and swizzled tile addressing, with
    wave-local steps and with a workgroup barrier after every step, on overlapping
    4-leg blocks and on mixed leg counts, against the oracle."""
    for name in ("k4 x3 overlapping", "k1-k5 mixed"):
        dsl, data = qxtest.gate_sequence_program(16, GATE_RUNS[name], seed=len(name))
        ex = Executor(dsl, data, precision=precision, batch=1, chain_variant=variant)
        out = np.asarray(ex.amplitudes([""])).reshape(-1)
        ref = O.amplitudes(dsl, data, [""])[0].reshape(-1, order="F")
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol, name
        act = [b for b in ex.stats()["batches"] if b["active"]][0]
        chains = [s for s in act["steps"] if s["kernel"] == "chain"]
        parts = [s["config"].split("/") for s in chains]
        assert chains and all(("fma", "mfma", "mfma3")[variant & 3] in p for p in parts)
        assert all(("swz" in p) == bool(variant & 4) for p in parts)
        if variant & 8:
            assert not any("wl" in p for p in parts)
