"""Output sampling: Philox pins and the rejection oracle on the CPU; the device
sampler (qx_amplitudes_uniform, qx_sample_rejection) against the oracle on the GPU.

Device results are compared with the oracle exactly (same candidates, same
accept/reject decisions, same order) and amplitudes within 1e-12 relative; the
sampled distribution is checked against exact probabilities."""

import math

import numpy as np
import pytest

import qxoracle as O
import qxsampling as S
import qxtest
from qxhip import Executor, circuits, simulation


# ---------------------------------------------------------------------- CPU pins
def test_philox_known_answer_vector():
    """Random123 known-answer vector for philox4x64_10, counter 0 and key 0."""
    assert S.philox4x64((0, 0, 0, 0), (0, 0)) == [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC,
                                                   0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]


@pytest.mark.parametrize("ctr,key", [((0, 0, 0, 0), (0, 0)), ((5, 0, 0, 0), (7, 9)),
                                     ((2 ** 64 - 2, 3, 0, 1), (2 ** 63, 12345))])
def test_philox_matches_numpy_bit_generator(ctr, key):
    """numpy's Philox is the same 4x64-10 generator; it steps the counter once
    before producing its first block."""
    g = np.random.Philox(counter=np.array(ctr, dtype=np.uint64), key=np.array(key, dtype=np.uint64))
    raw = [int(x) for x in g.random_raw(4)]
    c0 = (ctr[0] + 1) % 2 ** 64
    carry = 1 if c0 == 0 else 0
    nxt = (c0, (ctr[1] + carry) % 2 ** 64, ctr[2], ctr[3])
    assert raw == S.philox4x64(nxt, key)


def test_candidate_bits_and_uniforms():
    w = S.philox4x64((3, 0, 0, 0), (11, 0))
    b = S.candidate(3, 70, 11)
    assert len(b) == 70 and set(b) <= {"0", "1"}
    assert [int(c) for c in b[:64]] == [(w[0] >> q) & 1 for q in range(64)]
    assert [int(c) for c in b[64:]] == [(w[1] >> q) & 1 for q in range(6)]
    us = [S.uniform_u(j, 5) for j in range(2000)]
    assert all(0 <= u < 1 for u in us) and abs(np.mean(us) - 0.5) < 0.03
    bits = S.uniform_bitstrings(20, 500, 1)
    ones = np.array([[int(c) for c in s] for s in bits]).mean(0)
    assert np.all(np.abs(ones - 0.5) < 0.1)


def test_rejection_oracle_hand_worked():
    """Fixed M: accept exactly when u_j * M < r_j; running-max M otherwise."""
    n, seed = 3, 17
    cands = S.uniform_bitstrings(n, 64, seed)
    table = {b: complex(0.1 + 0.05 * int(b, 2), 0.02) for b in set(cands)}
    amp_of = lambda bs: np.array([table[b] for b in bs])
    bits, amps, M, used = S.rejection(amp_of, n, 5, seed, 0.7, True, 64)
    exp, j = [], 0
    while len(exp) < 5:
        a = table[cands[j]]
        if S.uniform_u(j, seed) * 0.7 < abs(a) ** 2 * 8:
            exp.append(cands[j])
        j += 1
    assert bits == exp and used == j and M == 0.7
    bits, amps, M, used = S.rejection(amp_of, n, 5, seed, 1e-9, False, 64)
    rs = [(table[c].real ** 2 + table[c].imag ** 2) * 8 for c in cands[:used]]
    assert M == max(rs) and bits[0] == cands[0]     # the first candidate sets M and is accepted


def test_rejection_oracle_distribution_is_exact_with_fixed_bound():
    """With M >= max r, accepted samples follow |amp|^2 exactly (chi-square)."""
    circ = circuits.create_rqc_circuit(2, 2, 6, seed=3)
    psi = O.statevector(4, qxtest.circuit_gates(circ))
    amp_of = lambda bs: O.sv_amplitudes(psi, bs)
    p = np.abs(psi.reshape(-1)) ** 2
    M = float(p.max() * 16) * 1.0001
    bits, _, _, _ = S.rejection(amp_of, 4, 4000, 99, M, True, 10 ** 6)
    probs = np.abs(amp_of(list(simulation.amplitudes_all(4)))) ** 2
    idx = {b: i for i, b in enumerate(simulation.amplitudes_all(4))}
    counts = np.bincount([idx[b] for b in bits], minlength=16)
    exp = probs * len(bits)
    chi2 = float(((counts - exp) ** 2 / np.maximum(exp, 1e-12)).sum())
    assert chi2 < 45.0          # 15 degrees of freedom: p ~ 1e-4


# ---------------------------------------------------------------------- GPU
def _executor(rows=3, cols=3, depth=8, nslice=0, batch=64):
    circ = circuits.create_rqc_circuit(rows, cols, depth, seed=5)
    t, plan, g = qxtest.graph_for(circ, planner="greedy" if nslice else "sweep", nslice=nslice)
    return circ, g, Executor(g.to_dsl(), g.data, batch=batch)


@pytest.mark.gpu
@pytest.mark.parametrize("nslice", [0, 2])
def test_device_uniform_matches_oracle(nslice):
    circ, g, ex = _executor(nslice=nslice)
    bits, amps = ex.amplitudes_uniform(100, 2024)
    assert bits == S.uniform_bitstrings(9, 100, 2024)
    ref = O.amplitudes(g.to_dsl(), g.data, bits)
    assert np.max(qxtest.rel_err(amps, ref)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("fix_M,M", [(False, 1e-4), (True, 3.0), (False, 2.0)])
def test_device_rejection_matches_oracle(fix_M, M):
    circ, g, ex = _executor(batch=32)
    dsl = g.to_dsl()
    amp_of = lambda bs: O.amplitudes(dsl, g.data, bs)
    bits, amps, m_out, used = ex.sample_rejection(40, 77, M, fix_M, 5000)
    rb, ra, rm, ru = S.rejection(amp_of, 9, 40, 77, M, fix_M, 5000)
    assert bits == rb and used == ru
    assert abs(m_out - rm) <= 1e-12 * rm
    assert np.max(qxtest.rel_err(amps, ra)) <= 1e-12


@pytest.mark.gpu
def test_device_rejection_distribution_and_limits():
    circ, g, ex = _executor(2, 2, 6, batch=256)
    psi = O.statevector(4, qxtest.circuit_gates(circ))
    p = np.abs(psi.reshape(-1)) ** 2
    bits, amps, m_out, used = ex.sample_rejection(3000, 5, float(p.max() * 16) * 1.0001, True, 10 ** 6)
    idx = {b: i for i, b in enumerate(simulation.amplitudes_all(4))}
    counts = np.bincount([idx[b] for b in bits], minlength=16)
    probs = np.abs(O.sv_amplitudes(psi, list(simulation.amplitudes_all(4)))) ** 2
    exp = probs * len(bits)
    assert float(((counts - exp) ** 2 / np.maximum(exp, 1e-12)).sum()) < 45.0
    from qxhip._native import QxError
    with pytest.raises(QxError):                       # runs out of candidates
        ex.sample_rejection(100, 5, 1e6, True, 50)
    out = simulation.run_outputs(ex, simulation.output_params_dict(4, 7, "Uniform", seed=3))
    assert list(out) == list(dict.fromkeys(S.uniform_bitstrings(4, 7, 3)))   # repeats collapse
    lst = simulation.run_outputs(ex, simulation.output_params_dict(4, bitstrings=["0101", "1111"]))
    assert np.allclose([lst["0101"], lst["1111"]], O.sv_amplitudes(psi, ["0101", "1111"]), atol=1e-14)
