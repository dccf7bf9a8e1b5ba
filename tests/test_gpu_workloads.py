"""GPU parity on the saved benchmark plans at their full sizes.

5x5 grid, depth 20 (25 qubits, complex128, no slicing, 256 seeded bitstrings; the
plan `bench.py` times): all 256 amplitudes against the FP64 statevector (2^25
entries, committed in `tests/golden/rqc_5x5_d20_amplitudes.npz` by
`tests/golden/make_rqc_5x5_d20.py`) within 1e-10 relative per amplitude, and 8 of
them against the oracle's contraction of the same plan within 1e-12 -- with the
heuristic kernel choice, after the autotuner has picked its kernels, and when the
tuned call is replayed from a captured graph."""

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
