"""Circuit -> tensor network conversion (the upstream end of the contraction path).

Mirrors `convert_to_tnc` (`src/tn_conversion.jl:10-40`) and the QXTns calls it
makes (`TensorNetworkCircuit`, `push!`, `add_input!`, `add_output!`).  QXTns is
not available, so its semantics are restated from what the reference's tests and
documents show:

* gate tensors carry legs `[out_q for q in qubits] + [in_q for non-diagonal q]`;
  a gate that is diagonal in qubit q keeps that qubit's wire index as a single
  shared ("hyper") leg instead of an in/out pair
  (`test/test_contraction_planning.jl:39-44`, `docs/src/features.md:5-10`);
* with `decompose=True` a two-qubit gate is split by SVD into two tensors joined
  by a bond, the 3-tensor GHZ circuit becoming 5 tensors
  (`test/test_tn_conversion.jl:16-18`; DSL example `docs/src/users_guide.md:71-90`);
* input and output tensors are rank-1 vectors on the first and last wire index of
  each qubit; bitstring characters '0', '1', '+', '-' select <0|, <1|, <+|, <-|
  (`docs/src/basics.md:55-63`).

Tensor data is a numpy array whose axis j is leg j; flattening in Julia's
column-major order (`reshape(data, prod(dim))`, `src/compute_graph/tensor_cache.jl:53`)
is `array.reshape(-1, order="F")`.
"""

from __future__ import annotations

import copy
import math
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from .circuits import Circuit, gate_matrix, gate_qubits

_S2 = 1.0 / math.sqrt(2.0)
OUTPUT_VECTORS = {
    "0": np.array([1.0, 0.0], dtype=np.complex128),
    "1": np.array([0.0, 1.0], dtype=np.complex128),
    "+": np.array([_S2, _S2], dtype=np.complex128),
    "-": np.array([_S2, -_S2], dtype=np.complex128),
}


def basis_vector(ch: str) -> np.ndarray:
    try:
        return OUTPUT_VECTORS[ch].copy()
    except KeyError:
        raise ValueError(f"bitstring character {ch!r} is not one of 0, 1, +, -") from None


@dataclass
class Tensor:
    indices: list          # index ids, one per leg (leg j = axis j of data)
    data: np.ndarray

    def copy(self) -> "Tensor":
        return Tensor(list(self.indices), self.data.copy())


class TensorNetwork:
    """Tensors keyed by name; an index shared by several tensors is a (hyper)edge."""

    def __init__(self):
        self.tensors: "OrderedDict[str, Tensor]" = OrderedDict()
        self.dims: dict = {}
        self._next_index = 1
        self._next_tensor = 1

    def new_index(self, dim: int = 2) -> int:
        i = self._next_index
        self._next_index += 1
        self.dims[i] = dim
        return i

    def next_tensor_id(self) -> str:
        s = f"t{self._next_tensor}"
        self._next_tensor += 1
        return s

    def push(self, indices: list, data: np.ndarray, name: str | None = None) -> str:
        if data.ndim != len(indices):
            raise ValueError("tensor rank does not match its index list")
        for i, d in zip(indices, data.shape):
            if self.dims.get(i, d) != d:
                raise ValueError(f"index {i} has dim {self.dims[i]}, tensor leg has {d}")
            self.dims[i] = d
        name = name or self.next_tensor_id()
        self.tensors[name] = Tensor(list(indices), np.asarray(data, dtype=np.complex128))
        return name

    def holders(self, index: int) -> list:
        return [n for n, t in self.tensors.items() if index in t.indices]

    def index_holders(self) -> dict:
        out: dict = {}
        for n, t in self.tensors.items():
            for i in t.indices:
                out.setdefault(i, []).append(n)
        return out

    def __len__(self) -> int:
        return len(self.tensors)

    def __getitem__(self, name: str) -> Tensor:
        return self.tensors[name]

    def __iter__(self):
        return iter(self.tensors)

    def copy(self) -> "TensorNetwork":
        return copy.deepcopy(self)


def _diagonal_in(t: np.ndarray, ax_out: int, ax_in: int) -> bool:
    """True if t vanishes whenever the two axes take different values."""
    d = t.shape[ax_out]
    for a in range(d):
        for b in range(d):
            if a != b and np.any(np.take(np.take(t, [a], axis=ax_out), [b], axis=ax_in)):
                return False
    return True


def _take_diagonal(t: np.ndarray, ax_keep: int, ax_drop: int) -> np.ndarray:
    """Collapse two equal-dim axes to their diagonal, keeping it at ax_keep."""
    d = t.shape[ax_keep]
    parts = [np.take(np.take(t, [a], axis=ax_keep), [a], axis=ax_drop) for a in range(d)]
    out = np.concatenate(parts, axis=ax_keep)
    return np.squeeze(out, axis=ax_drop)


class TensorNetworkCircuit:
    """A tensor network plus the wire bookkeeping of a circuit (QXTns'
    `TensorNetworkCircuit`, re-exported at `src/QXTools.jl:8`)."""

    def __init__(self, num_qubits: int):
        self.tn = TensorNetwork()
        self.num_qubits = num_qubits
        self.input_indices = [self.tn.new_index(2) for _ in range(num_qubits)]
        self.output_indices = list(self.input_indices)
        self.input_tensors: list = []
        self.output_tensors: list = []
        self.output_qubits: list = []  # qubit of each output tensor
        self.gate_tensors: list = []   # per pushed gate: (tensor names, qubits), time order

    def qubits(self) -> int:
        return self.num_qubits

    def open_indices(self) -> list:
        """Wire ends not capped by an input or output tensor: open legs of the
        network even when a diagonal gate makes them hyperedges."""
        out = []
        if not self.input_tensors:
            out += self.input_indices
        capped = set(self.output_qubits)
        out += [i for q, i in enumerate(self.output_indices, start=1) if q not in capped and i not in out]
        return out

    def __len__(self) -> int:
        return len(self.tn)

    def keys(self):
        return list(self.tn.tensors.keys())

    def copy(self) -> "TensorNetworkCircuit":
        return copy.deepcopy(self)

    def tensor_data(self, name: str) -> np.ndarray:
        return self.tn[name].data

    # -- gates -------------------------------------------------------------------
    def push(self, qubits: list, mat: np.ndarray, decompose: bool = False) -> list:
        """Append a gate (QXTns `push!(tnc, qubits, mat; decompose)`, called at
        `src/tn_conversion.jl:15`).  Returns the names of the tensors added."""
        k = len(qubits)
        if mat.shape != (2 ** k, 2 ** k):
            raise ValueError("gate matrix shape does not match qubit count")
        # axes [o_1..o_k, i_1..i_k]; first qubit fastest (column-major reshape)
        t = np.asarray(mat, dtype=np.complex128).reshape((2,) * (2 * k), order="F")
        if decompose and k == 2:
            return self._push_decomposed(qubits, t)
        legs_out, legs_in, keep = [], [], []
        diag = [_diagonal_in(t, j, k + j) for j in range(k)]
        data = t
        # collapse diagonal qubits (highest axes first so positions stay valid)
        for j in reversed(range(k)):
            if diag[j]:
                data = _take_diagonal(data, j, k + j)
        for j, q in enumerate(qubits):
            w = self.output_indices[q - 1]
            if diag[j]:
                legs_out.append(w)
            else:
                o = self.tn.new_index(2)
                legs_out.append(o)
                legs_in.append(w)
                keep.append((q, o))
        name = self.tn.push(legs_out + legs_in, data)
        for q, o in keep:
            self.output_indices[q - 1] = o
        self.gate_tensors.append(([name], list(qubits)))
        return [name]

    def _push_decomposed(self, qubits: list, t: np.ndarray) -> list:
        """Split a two-qubit gate [o1,o2,i1,i2] by SVD into factors [o1,i1,bond] and
        [o2,i2,bond] (sqrt of the singular values to each side).  Each factor that is
        diagonal in (out, in) keeps its wire as one hyper leg; a factor that is then
        also diagonal between that wire and the bond is a copy tensor, so the bond is
        identified with the wire: the factor becomes a rank-1 tensor on the wire and
        the other factor carries the wire in place of the bond.  The CX control side
        is such a factor, which gives the hyperedge census of the reference's test
        (test/test_contraction_planning.jl:39-44)."""
        m = t.transpose(0, 2, 1, 3).reshape(4, 4, order="F")
        u, s, vh = np.linalg.svd(m)
        r = int(np.sum(s > 1e-12 * s[0]))
        sq = np.sqrt(s[:r])
        fa = (u[:, :r] * sq).reshape(2, 2, r, order="F")             # [o1, i1, bond]
        fb = (sq[:, None] * vh[:r, :]).T.reshape(2, 2, r, order="F")   # [o2, i2, bond]
        parts = []
        for q, f in ((qubits[0], fa), (qubits[1], fb)):
            w = self.output_indices[q - 1]
            if _diagonal_in(f, 0, 1):
                parts.append((q, _take_diagonal(f, 0, 1), [w], None))     # [w, bond]
            else:
                parts.append((q, f, [None, w], w))                        # [o, w, bond]
        bond = None
        for j, (q, f, legs, _) in enumerate(parts):
            if legs == [self.output_indices[q - 1]] and f.shape[1] == f.shape[0] and _diagonal_in(f, 0, 1):
                bond = legs[0]
                parts[j] = (q, _take_diagonal(f, 0, 1), legs, "copy")    # rank-1 on the wire
                break
        if bond is None:
            bond = self.new_index(r)
        names = []
        for q, f, legs, kind in parts:
            if kind == "copy":
                names.append(self.tn.push(legs, f))
                continue
            if legs[0] is None:
                o = self.new_index(2)
                names.append(self.tn.push([o, legs[1], bond], f))
                self.output_indices[q - 1] = o
            else:
                names.append(self.tn.push(legs + [bond], f))
        self.gate_tensors.append((names, list(qubits)))
        return names

    def new_index(self, dim: int) -> int:
        return self.tn.new_index(dim)

    # -- boundary tensors ----------------------------------------------------------
    def add_input(self, bits: str | None = None) -> None:
        """Rank-1 input tensors on each qubit's first wire index (QXTns `add_input!`,
        `src/tn_conversion.jl:37`).  Default state |0...0>."""
        bits = bits or "0" * self.num_qubits
        if len(bits) != self.num_qubits:
            raise ValueError("input bitstring length differs from qubit count")
        if self.input_tensors:
            for name, ch in zip(self.input_tensors, bits):
                self.tn[name].data = basis_vector(ch)
            return
        for q, ch in enumerate(bits):
            self.input_tensors.append(self.tn.push([self.input_indices[q]], basis_vector(ch)))

    def add_output(self, bits: str | None = None, open_qubits=()) -> None:
        """Rank-1 output tensors <b_q| on each qubit's last wire index (QXTns
        `add_output!`, `src/tn_conversion.jl:38`, `src/simulation.jl:88`).  When the
        outputs exist already their values are replaced.  Qubits in `open_qubits` get
        no output tensor: their last wire stays an open leg of the network."""
        bits = bits or "0" * self.num_qubits
        if len(bits) != self.num_qubits:
            raise ValueError("output bitstring length differs from qubit count")
        if self.output_tensors:
            for name, q in zip(self.output_tensors, self.output_qubits):
                self.tn[name].data = basis_vector(bits[q - 1])
            return
        skip = set(open_qubits)
        for q, ch in enumerate(bits, start=1):
            if q in skip:
                continue
            self.output_tensors.append(self.tn.push([self.output_indices[q - 1]], basis_vector(ch)))
            self.output_qubits.append(q)


def convert_to_tnc(circ: Circuit, input: str | None = None, output: str | None = None,
                   no_input: bool = False, no_output: bool = False,
                   decompose: bool = False, open_outputs=()) -> TensorNetworkCircuit:
    """Circuit -> TensorNetworkCircuit (`src/tn_conversion.jl:30-40`).  `open_outputs`
    leaves those qubits' output legs open (one contraction then yields 2^m
    amplitudes)."""
    tnc = TensorNetworkCircuit(circ.num_qubits)
    for g in circ.ops:
        tnc.push(gate_qubits(g), gate_matrix(g), decompose=decompose)
    if not no_input:
        tnc.add_input(input)
    if not no_output:
        tnc.add_output(output, open_qubits=open_outputs)
    return tnc


def gate_tensor_map(circ: Circuit, decompose: bool = False) -> tuple:
    """Convert `circ` and also return, per gate index, the tensor names it produced
    (used to map tensor gradients back to gate parameters)."""
    tnc = TensorNetworkCircuit(circ.num_qubits)
    names = []
    for g in circ.ops:
        names.append(tnc.push(gate_qubits(g), gate_matrix(g), decompose=decompose))
    return tnc, names
