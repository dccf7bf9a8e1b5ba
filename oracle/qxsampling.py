"""CPU oracle for output sampling.  TEST INFRASTRUCTURE (same rules as qxoracle:
only tests/, smoke() and bench's cpu_baseline may import it).

Restates the output methods of `output_params_dict` (`src/outputs.jl:47-78`) as the
library runs them on the device:

* Candidate bitstrings come from Philox4x64-10 (Salmon et al., SC 2011): candidate
  j of stream `seed` has character q = bit q of the block at counter (j,0,0,0),
  key (seed,0).  Julia's MersenneTwister stream of `amplitudes_uniform`
  (`src/simulation.jl:24-28`) cannot be reproduced here; this generator is the
  build's own, pinned below.
* :Rejection is empirical-supremum rejection sampling as `docs/src/features.md:
  73-90` describes it (QXContexts' implementation is not available): uniform
  proposal, r_j = |a_j|^2 2^n, M_j = M_{j-1} (fix_M) or max(M_{j-1}, r_j), accept
  when u_j * M_j < r_j, u_j = top 53 bits of the block at key (seed,1) * 2^-53.

Pins (tests/test_sampling.py): the Random123 known-answer vector of
philox4x64_10 (counter 0, key 0), numpy's own Philox bit generator (which steps
its counter before the first block), and, for rejection, exactness of the
sampled distribution when M is fixed at or above max r (chi-square on a small
circuit) and hand-worked accept/reject sequences.
"""

from __future__ import annotations

import numpy as np

_M0, _M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
_W0, _W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
_MASK = (1 << 64) - 1


def philox4x64(counter, key) -> list:
    """Philox4x64 with 10 rounds on Python integers."""
    c = [int(x) & _MASK for x in counter]
    k0, k1 = int(key[0]) & _MASK, int(key[1]) & _MASK
    for _ in range(10):
        p0, p1 = _M0 * c[0], _M1 * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k0, p1 & _MASK, (p0 >> 64) ^ c[3] ^ k1, p0 & _MASK]
        k0, k1 = (k0 + _W0) & _MASK, (k1 + _W1) & _MASK
    return c


def candidate(j: int, n: int, seed: int) -> str:
    w = philox4x64((j, 0, 0, 0), (seed, 0))
    return "".join("01"[(w[q >> 6] >> (q & 63)) & 1] for q in range(n))


def uniform_u(j: int, seed: int) -> float:
    return (philox4x64((j, 0, 0, 0), (seed, 1))[0] >> 11) * 2.0 ** -53


def uniform_bitstrings(n: int, num: int, seed: int, first: int = 0) -> list:
    return [candidate(first + j, n, seed) for j in range(num)]


def rejection(amp_of, n: int, num_samples: int, seed: int, M: float, fix_M: bool, max_candidates: int):
    """(accepted bitstrings, their amplitudes, final M, candidates consumed).
    `amp_of(list of bitstrings) -> complex array`."""
    out, amps = [], []
    j = 0
    scale = float(2.0 ** n)
    chunk = 256
    while len(out) < num_samples and j < max_candidates:
        cands = uniform_bitstrings(n, min(chunk, max_candidates - j), seed, j)
        a = np.asarray(amp_of(cands), dtype=np.complex128).reshape(-1)
        for b, x in zip(cands, a):
            r = (x.real * x.real + x.imag * x.imag) * scale
            if not fix_M:
                M = max(M, r)
            u = uniform_u(j, seed)
            j += 1
            if u * M < r:
                out.append(b)
                amps.append(x)
                if len(out) == num_samples:
                    break
    return out, np.asarray(amps, dtype=np.complex128), M, j
