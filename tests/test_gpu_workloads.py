"""GPU parity on the saved benchmark plans at their full sizes.

5x5 grid, depth 20 (25 qubits, complex128, no slicing, 256 seeded bitstrings; the
plan `bench.py` times): all 256 amplitudes against the FP64 statevector (2^25
entries, committed in `tests/golden/rqc_5x5_d20_amplitudes.npz` by
`tests/golden/make_rqc_5x5_d20.py`) within 1e-10 relative per amplitude, and 8 of
them against the oracle's contraction of the same plan within 1e-12 -- with the
heuristic kernel choice, after the autotuner has picked its kernels, and when the
tuned call is replayed from a captured graph.

The other three workloads at their full sizes, against the oracle's contraction of
the same plans (`tests/golden/make_configs_3_4_5.py`; their statevectors are out of
reach): the sliced 6x6 and 7x7 plans with slices fused and looped, and the
parameterised 6x6 plan of the gradient benchmark, whose gradient is also checked
against central differences."""

import hashlib
import os

import numpy as np
import pytest

from qxhip import Executor, workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rqc_5x5_d20_amplitudes.npz")


def _rel(a, r):
    return np.abs(np.asarray(a) - r) / np.abs(r)


@pytest.fixture(scope="module")
def config2():
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    with np.load(GOLDEN, allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    assert list(gold["bits"]) == bits
    return dsl, data, bits, gold


@pytest.mark.parametrize("autotune,graph", [(0, 0), (1, 1)])
def test_rqc_5x5_d20_full_size_c128(config2, autotune, graph):
    dsl, data, bits, gold = config2
    ex = Executor(dsl, data, precision="c128", batch=256, autotune=autotune, graph=graph)
    try:
        first = ex.amplitudes(bits)        # the tuning call when autotune is on
        out = ex.amplitudes(bits)          # tuned kernels (captured, then replayed below)
        again = ex.amplitudes(bits)
    finally:
        ex.close()
    for a in (first, out, again):
        assert _rel(a, gold["sv_amps"]).max() <= 1e-10
        assert _rel(a[gold["oracle_idx"]], gold["oracle_amps"]).max() <= 1e-12
    assert np.array_equal(out, again)


def test_rqc_5x5_d20_c64_against_c128(config2):
    """complex64 storage with FP32 accumulation on the same plan: normwise 2e-6 and
    per amplitude 1e-4 against the statevector (FP32 sums of long contractions)."""
    dsl, data, bits, gold = config2
    ex = Executor(dsl, data, precision="c64_4m", batch=256)
    try:
        out = ex.amplitudes(bits)
    finally:
        ex.close()
    ref = gold["sv_amps"]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) <= 2e-6
    assert _rel(out, ref).max() <= 1e-4


def _oracle_fixture(name):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"{name}_oracle.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _nerr(a, r):
    return float(np.linalg.norm(np.asarray(a).reshape(-1) - r.reshape(-1)) / np.linalg.norm(r))


@pytest.mark.parametrize("name", ["rqc_6x6_d12_s6", "rqc_7x7_d12_s10_open5"])
def test_sliced_workloads_against_oracle(name):
    """The 6x6 depth-12 workload (2^6 sliced bonds, 16 of its bitstrings) and the 7x7
    depth-12 workload (2^10 sliced bonds, 5 open outputs, 2 bitstring prefixes of 32
    amplitudes each) at their full size against the oracle's contraction of the same
    plans (tests/golden/make_configs_3_4_5.py): complex128 within 1e-10 per amplitude
    with all slices fused inline and looped, and the workload's own complex64 storage
    within 2e-5 normwise."""
    w = workloads.WORKLOADS[name]
    dsl, data, _ = w.load()
    gold = _oracle_fixture(name)
    assert hashlib.sha256(dsl.encode()).hexdigest() == str(gold["dsl_sha256"])
    bits = [str(b) for b in gold["bits"]]
    ref = gold["amps"]
    for fuse in (1, 0):
        ex = Executor(dsl, data, precision="c128", batch=len(bits), fuse_slices=fuse, batch_last=w.batch_last)
        try:
            out = np.asarray(ex.amplitudes(bits)).reshape(len(bits), -1)
        finally:
            ex.close()
        assert _rel(out, ref).max() <= 1e-10, (fuse, _rel(out, ref).max())
    ex = Executor(dsl, data, precision=w.precision, batch=len(bits), batch_last=w.batch_last)
    try:
        out = np.asarray(ex.amplitudes(bits)).reshape(len(bits), -1)
    finally:
        ex.close()
    assert _nerr(out, ref) <= 2e-5


def test_param_workload_against_oracle_and_finite_differences():
    """The parameterised 6x6 depth-6 workload (648 angles, 2^4 sliced bonds, 64
    bitstrings) as the gradient benchmark plans it: amplitudes within 1e-10 of the
    oracle's contraction of the same plan, and the gradient of sum |amp|^2 against
    central differences of the GPU's own loss on three angles (h = 1e-5)."""
    from qxhip import gradients
    w = workloads.PARAM_WORKLOAD
    gold = _oracle_fixture("param_6x6_d6_s4")
    pc = gradients.ParamCircuit(w.circuit(), planner="greedy", nslice=w.sliced, precision="c128",
                                batch=64, **w.plan_kwargs)
    assert hashlib.sha256(pc.dsl.encode()).hexdigest() == str(gold["dsl_sha256"])
    bits = [str(b) for b in gold["bits"]]
    amps = np.asarray(pc.amplitudes(bits)).reshape(len(bits), -1)
    assert _rel(amps, gold["amps"]).max() <= 1e-10
    loss, grad, _ = pc.loss_and_grad(bits)
    assert loss == pytest.approx(float(np.sum(np.abs(gold["amps"]) ** 2)), rel=1e-10)
    th = pc.params().copy()
    h = 1e-5
    for p in (0, len(th) // 2, len(th) - 1):
        e = np.zeros_like(th)
        e[p] = h
        pc.set_params(th + e)
        lp = float(np.sum(np.abs(pc.amplitudes(bits)) ** 2))
        pc.set_params(th - e)
        lm = float(np.sum(np.abs(pc.amplitudes(bits)) ** 2))
        pc.set_params(th)
        fd = (lp - lm) / (2 * h)
        assert abs(grad[p] - fd) <= 1e-6 * np.max(np.abs(grad)), (p, grad[p], fd)
