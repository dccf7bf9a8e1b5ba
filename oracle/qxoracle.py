"""CPU FP64 oracle for the tensor-network contraction path.  TEST INFRASTRUCTURE.

Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may import this module, and only as the checker / CPU baseline.  The product path
(`libqxhip.so` and the `qxhip` package) never calls it.

It restates, in plain numpy, what the reference's executor does with a saved plan
(`.qx` DSL + initial tensors), so it reads the DSL on its own (it does not share
the product's parser):

* DSL semantics: `docs/src/users_guide.md:37-170` (load / output / view / ncon /
  save; label 0 = scalar; batched "hyper" labels appear in A, B and C).
* Views keep the rank, the sliced leg becoming size 1 (`users_guide.md:107-122`,
  built at `src/compute_graph/compute_graph.jl:39-58`); chained views compose.
* Output tensors take <b_q| from character q of the bitstring, '0','1','+','-'
  (`docs/src/basics.md:55-63`; driver `src/simulation.jl:86-91`).
* Each ncon is evaluated TTGT-style as TensorOperations does it (the arithmetic
  behind QXTns `contract_pair!`, `Project.toml:38`): transpose A to (H,M,K) and B to
  (H,K,N), one batched complex128 matmul (OpenBLAS zgemm), reshape and transpose to
  C's label order.
* The slice sum runs over all assignments of v1..vk with v1 fastest (Julia
  CartesianIndices order) and is Neumaier-compensated.

Pins (see tests/test_oracle.py): GHZ-3/GHZ-5 closed forms from the reference's
tests (`test/test_tn_conversion.jl:38-40`, `test/test_simulation.jl:24-25`), an
independent FP64 statevector simulator (below), `np.einsum` on random steps,
Sum |amp|^2 = 1, and the slice-sum identity.
"""

from __future__ import annotations

import itertools
import math

import numpy as np

_S2 = 1.0 / math.sqrt(2.0)
_OUT = {"0": (1.0, 0.0), "1": (0.0, 1.0), "+": (_S2, _S2), "-": (_S2, -_S2)}


# ----------------------------------------------------------------------------- DSL
def parse(text: str) -> list:
    """DSL text -> list of (op, fields) tuples (`users_guide.md:95-165`)."""
    out = []
    for raw in text.splitlines():
        s = raw.strip()
        if not s or s.startswith("#"):
            continue
        f = s.split()
        out.append((f[0], f[1:]))
    return out


def _labels(s: str) -> list:
    return [] if s == "0" else [int(x) for x in s.split(",")]


def slice_vars(cmds: list) -> list:
    """[(name, dim)] of slice variables in order of first appearance, sorted v1..vk."""
    seen = {}
    for op, f in cmds:
        if op == "view":
            seen.setdefault(f[2], int(f[4]))
    return sorted(seen.items(), key=lambda kv: int(kv[0][1:]) if kv[0][1:].isdigit() else kv[0])


def slice_assignments(cmds: list):
    """All assignments {var: 0-based value}, first variable fastest."""
    vs = slice_vars(cmds)
    ranges = [range(d) for _, d in vs]
    for combo in itertools.product(*reversed(ranges)):
        vals = list(reversed(combo))
        yield {name: v for (name, _), v in zip(vs, vals)}


def output_qubits(cmds: list) -> list:
    return [int(f[1]) for op, f in cmds if op == "output"]


# ----------------------------------------------------------------------------- ncon
def ncon(A: np.ndarray, al: list, B: np.ndarray, bl: list, cl: list) -> np.ndarray:
    """C[cl] = sum over (al & bl) - cl of A[al] * B[bl], TTGT with one matmul."""
    for labs in (al, bl, cl):
        if len(set(labs)) != len(labs):
            raise ValueError("repeated label within one tensor")
    if A.ndim != len(al) or B.ndim != len(bl):
        raise ValueError("label count differs from tensor rank")
    Hl = [l for l in al if l in bl and l in cl]
    Kl = [l for l in al if l in bl and l not in cl]
    Ml = [l for l in al if l not in bl]
    Nl = [l for l in bl if l not in al]
    if any(l not in cl for l in Ml + Nl) or set(cl) != set(Hl + Ml + Nl):
        raise ValueError("label classes are not H/K/M/N")
    da = dict(zip(al, A.shape))
    db = dict(zip(bl, B.shape))
    for l in Hl + Kl:
        if da[l] != db[l]:
            raise ValueError(f"label {l}: dims {da[l]} vs {db[l]}")
    p = lambda ls, d: int(np.prod([d[l] for l in ls], dtype=np.int64))
    h, m, k, n = p(Hl, da), p(Ml, da), p(Kl, da), p(Nl, db)
    At = A.transpose([al.index(l) for l in Hl + Ml + Kl]).reshape(h, m, k)
    Bt = B.transpose([bl.index(l) for l in Hl + Kl + Nl]).reshape(h, k, n)
    Ct = np.matmul(At, Bt)
    order = Hl + Ml + Nl
    shape = [da[l] for l in Hl + Ml] + [db[l] for l in Nl]
    C = Ct.reshape(shape)
    return C.transpose([order.index(l) for l in cl]) if cl else C.reshape(())


# ----------------------------------------------------------------------------- executor
def _fetch(data: dict, key: str, dims: tuple) -> np.ndarray:
    a = np.asarray(data[key], dtype=np.complex128)
    if a.ndim == 1 and len(dims) != 1:
        a = a.reshape(dims, order="F")
    if tuple(a.shape) != tuple(dims):
        raise ValueError(f"data {key} has shape {a.shape}, load expects {dims}")
    return a


def contract_slice(cmds: list, data: dict, bits: str, assign: dict, override: dict | None = None) -> np.ndarray:
    """Evaluate the DAG for one bitstring and one slice assignment.  `override`
    replaces the tensor of named load nodes (used for environment tensors)."""
    env: dict = {}
    result = None
    for op, f in cmds:
        if op == "load":
            dims = tuple(int(x) for x in f[2].split(",")) if len(f) > 2 else ()
            env[f[0]] = override[f[0]] if override and f[0] in override else _fetch(data, f[1], dims)
        elif op == "output":
            q, d = int(f[1]), int(f[2])
            if d != 2:
                raise ValueError("output tensors have dim 2")
            env[f[0]] = np.array(_OUT[bits[q - 1]], dtype=np.complex128)
        elif op == "view":
            new, old, var, pos = f[0], f[1], f[2], int(f[3])
            t = env[old]
            v = assign[var]
            idx = (slice(None),) * (pos - 1) + (slice(v, v + 1),)
            env[new] = t[idx]
        elif op == "ncon":
            c, cl, a, al, b, bl = f[0], _labels(f[1]), f[2], _labels(f[3]), f[4], _labels(f[5])
            env[c] = ncon(env[a], al, env[b], bl, cl)
        elif op == "save":
            result = env[f[1]]
        else:
            raise ValueError(f"unknown instruction {op}")
    return result


class Neumaier:
    """Compensated summation (Neumaier's variant of Kahan), real and imaginary
    parts carried separately."""

    def __init__(self, shape=()):
        self.s = np.zeros(tuple(shape) + (2,))
        self.c = np.zeros(tuple(shape) + (2,))

    def add(self, x):
        x = np.asarray(x)
        xv = np.stack([np.real(x), np.imag(x)], axis=-1).astype(np.float64)
        t = self.s + xv
        self.c += np.where(np.abs(self.s) >= np.abs(xv), (self.s - t) + xv, (xv - t) + self.s)
        self.s = t

    def value(self):
        v = self.s + self.c
        return v[..., 0] + 1j * v[..., 1]


def amplitudes(dsl: str, data: dict, bitstrings: list, slices=None) -> np.ndarray:
    """Amplitudes for each bitstring, summed over slices (all, or the given
    list of assignments).  Returns complex128 of shape (nb,) + final tensor shape."""
    cmds = parse(dsl)
    assigns = list(slice_assignments(cmds)) if slices is None else slices
    out = []
    for bits in bitstrings:
        acc = None
        for a in assigns:
            r = contract_slice(cmds, data, bits, a)
            if acc is None:
                acc = Neumaier(r.shape)
            acc.add(r)
        out.append(acc.value())
    return np.array(out)


def env_gradient(dsl: str, data: dict, bitstrings: list, amp_cot: np.ndarray) -> dict:
    """Reference gradient by environment tensors: the amplitude is linear in each
    load node's tensor, so d amp / d T_node[i] is the amplitude with that node's
    tensor replaced by the i-th basis tensor.  Returns, per data key,
    sum_b amp_cot_b conj(d amp_b / d T_key) summed over the nodes loading the key
    and over all slices.  Cost: one contraction per (node, element, bitstring,
    slice) -- small networks only."""
    cmds = parse(dsl)
    assigns = list(slice_assignments(cmds))
    cot = np.asarray(amp_cot, dtype=np.complex128).reshape(len(bitstrings), -1)
    out: dict = {}
    for op, f in cmds:
        if op != "load":
            continue
        name, key = f[0], f[1]
        dims = tuple(int(x) for x in f[2].split(",")) if len(f) > 2 else ()
        n = int(np.prod(dims, dtype=np.int64)) if dims else 1
        g = out.setdefault(key, np.zeros(n, dtype=np.complex128))
        for i in range(n):
            e = np.zeros(n, dtype=np.complex128)
            e[i] = 1.0
            ov = {name: e.reshape(dims, order="F") if dims else e.reshape(())}
            for b, bits in enumerate(bitstrings):
                d = sum(contract_slice(cmds, data, bits, a, ov) for a in assigns)
                g[i] += np.sum(cot[b] * np.conj(np.asarray(d).reshape(-1, order="F")))
    return out


def per_slice(dsl: str, data: dict, bits: str) -> list:
    cmds = parse(dsl)
    return [contract_slice(cmds, data, bits, a) for a in slice_assignments(cmds)]


# ----------------------------------------------------------------------------- statevector
def statevector(num_qubits: int, gates: list) -> np.ndarray:
    """FP64 statevector of U|0..0>, axis q-1 = qubit q.  `gates` is a list of
    (matrix 2^k x 2^k, [q1..qk]); the first listed qubit is the fastest index of
    the matrix (column-major reshape to [o1..ok, i1..ik])."""
    psi = np.zeros((2,) * num_qubits, dtype=np.complex128)
    psi[(0,) * num_qubits] = 1.0
    for mat, qs in gates:
        k = len(qs)
        t = np.asarray(mat, dtype=np.complex128).reshape((2,) * (2 * k), order="F")
        axes = [q - 1 for q in qs]
        psi = np.tensordot(t, psi, axes=(list(range(k, 2 * k)), axes))
        # tensordot puts the k output axes first; move them back to their qubits
        psi = np.moveaxis(psi, list(range(k)), axes)
    return psi


def sv_amplitude(psi: np.ndarray, bits: str) -> complex:
    v = psi
    for ch in bits:
        v = np.tensordot(np.array(_OUT[ch], dtype=np.complex128), v, axes=([0], [0]))
    return complex(v)


def sv_amplitudes(psi: np.ndarray, bitstrings: list) -> np.ndarray:
    return np.array([sv_amplitude(psi, b) for b in bitstrings])


# ----------------------------------------------------------------------------- stats
def step_shapes(dsl: str) -> list:
    """Per ncon step (B, M, N, K) products of H, M, N, K leg dims for one slice
    (sliced legs count 1), following the DAG's shapes."""
    cmds = parse(dsl)
    shapes: dict = {}
    out = []
    for op, f in cmds:
        if op == "load":
            shapes[f[0]] = tuple(int(x) for x in f[2].split(",")) if len(f) > 2 else ()
        elif op == "output":
            shapes[f[0]] = (int(f[2]),)
        elif op == "view":
            s = list(shapes[f[1]])
            s[int(f[3]) - 1] = 1
            shapes[f[0]] = tuple(s)
        elif op == "ncon":
            c, cl, a, al, b, bl = f[0], _labels(f[1]), f[2], _labels(f[3]), f[4], _labels(f[5])
            da, db = dict(zip(al, shapes[a])), dict(zip(bl, shapes[b]))
            pr = lambda ls, d: int(np.prod([d[l] for l in ls], dtype=np.int64))
            H = [l for l in al if l in bl and l in cl]
            K = [l for l in al if l in bl and l not in cl]
            M = [l for l in al if l not in bl]
            N = [l for l in bl if l not in al]
            out.append((pr(H, da), pr(M, da), pr(N, db), pr(K, da)))
            dd = {**db, **da}
            shapes[c] = tuple(dd[l] for l in cl)
    return out
