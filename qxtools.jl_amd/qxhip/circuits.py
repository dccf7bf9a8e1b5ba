"""Quantum circuits used as inputs of the contraction path.

Mirrors the circuit constructors of `src/circuits/circuits.jl:17-90`
(`create_test_circuit`, `create_ghz_circuit`, `create_rqc_circuit`,
`gate_matrix`, `gate_qubits`).  QXZoo (the package those functions wrap) is not
available, so the random-circuit generator here is this package's own seeded
construction:

* each cycle applies one single-qubit gate per qubit drawn from
  {sqrt(X), sqrt(Y), sqrt(W)} that differs from the previous one on that qubit,
* then CZ on one of four staggered nearest-neighbour edge patterns
  (horizontal even, vertical even, horizontal odd, vertical odd; cycle t uses
  pattern t mod 4).

Qubits are numbered from 1, as in `docs/src/basics.md:27`.  A two-qubit gate
matrix acts on qubits `[target, ctrl]`, the first listed qubit being the
fastest-varying index of the 4x4 matrix (`src/circuits/circuits.jl:83-90`,
`docs/src/basics.md:36-45`).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_S2 = 1.0 / math.sqrt(2.0)


def _sqrt_of_involution(g: np.ndarray) -> np.ndarray:
    """sqrt(G) = ((1+i) I + (1-i) G) / 2 for a Hermitian unitary G (G^2 = I)."""
    return 0.5 * ((1 + 1j) * np.eye(2) + (1 - 1j) * g)


I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = _S2 * np.array([[1, 1], [1, -1]], dtype=np.complex128)
W = _S2 * (X + Y)
SQRT_X = _sqrt_of_involution(X)
SQRT_Y = _sqrt_of_involution(Y)
SQRT_W = _sqrt_of_involution(W)
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
# [target, ctrl] ordering: basis index = b_target + 2*b_ctrl
CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)


def rz(theta: float) -> np.ndarray:
    return np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)]).astype(np.complex128)


def ry(theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def drz(theta: float) -> np.ndarray:
    """d Rz(theta) / d theta."""
    return np.diag([-0.5j * np.exp(-0.5j * theta), 0.5j * np.exp(0.5j * theta)]).astype(np.complex128)


def dry(theta: float) -> np.ndarray:
    """d Ry(theta) / d theta."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return 0.5 * np.array([[-s, -c], [c, -s]], dtype=np.complex128)


GATES = {
    "x": X, "y": Y, "z": Z, "h": H, "sx": SQRT_X, "sy": SQRT_Y, "sw": SQRT_W,
    "cz": CZ, "cx": CX,
}
_PARAM_GATES = {"rz": (rz, drz), "ry": (ry, dry)}


@dataclass
class Gate:
    """One gate call: `name`, qubits `[target, ctrl...]` (1-based), optional angle.

    `param` indexes the circuit's parameter vector for parameterised gates."""

    name: str
    qubits: tuple
    theta: float | None = None
    param: int | None = None


@dataclass
class Circuit:
    num_qubits: int
    ops: list = field(default_factory=list)

    def add(self, name: str, *qubits: int, theta=None, param=None) -> None:
        for q in qubits:
            if not 1 <= q <= self.num_qubits:
                raise ValueError(f"qubit {q} out of range 1..{self.num_qubits}")
        self.ops.append(Gate(name, tuple(qubits), theta, param))

    @property
    def num_params(self) -> int:
        return 1 + max((g.param for g in self.ops if g.param is not None), default=-1)

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params)
        for g in self.ops:
            if g.param is not None:
                out[g.param] = g.theta
        return out

    def with_params(self, theta: np.ndarray) -> "Circuit":
        c = Circuit(self.num_qubits)
        for g in self.ops:
            t = g.theta if g.param is None else float(theta[g.param])
            c.ops.append(Gate(g.name, g.qubits, t, g.param))
        return c


def gate_matrix(gate: Gate) -> np.ndarray:
    """Complex128 matrix of a gate call (`src/circuits/circuits.jl:74-76`)."""
    if gate.name in _PARAM_GATES:
        return _PARAM_GATES[gate.name][0](gate.theta)
    return GATES[gate.name].astype(np.complex128)


def gate_matrix_derivative(gate: Gate) -> np.ndarray:
    """d gate_matrix / d theta for a parameterised gate."""
    return _PARAM_GATES[gate.name][1](gate.theta)


def gate_qubits(gate: Gate) -> list:
    """Qubits a gate acts on, `[target, ctrl]` for two-qubit gates (`circuits.jl:83-90`)."""
    return list(gate.qubits)


def create_test_circuit() -> Circuit:
    """3-qubit GHZ preparation: H(1), CX(1,2), CX(2,3) (`circuits.jl:17-23`)."""
    c = Circuit(3)
    c.add("h", 1)
    c.add("cx", 2, 1)   # [target, ctrl]: control 1, target 2
    c.add("cx", 3, 2)
    return c


def create_ghz_circuit(n: int) -> Circuit:
    """n-qubit GHZ preparation (`circuits.jl:30-37`)."""
    c = Circuit(n)
    c.add("h", 1)
    for i in range(1, n):
        c.add("cx", i + 1, i)
    return c


def grid_edge_patterns(rows: int, cols: int) -> list:
    """The four staggered nearest-neighbour CZ patterns on a rows x cols grid.

    Qubit (r, c), 0-based, is number r*cols + c + 1."""
    q = lambda r, c: r * cols + c + 1
    hor_even = [(q(r, c), q(r, c + 1)) for r in range(rows) for c in range(0, cols - 1, 2)]
    hor_odd = [(q(r, c), q(r, c + 1)) for r in range(rows) for c in range(1, cols - 1, 2)]
    ver_even = [(q(r, c), q(r + 1, c)) for r in range(0, rows - 1, 2) for c in range(cols)]
    ver_odd = [(q(r, c), q(r + 1, c)) for r in range(1, rows - 1, 2) for c in range(cols)]
    return [hor_even, ver_even, hor_odd, ver_odd]


def create_rqc_circuit(rows: int, cols: int, depth: int, seed: int | None = None,
                       final_h: bool = False) -> Circuit:
    """Seeded random grid circuit (the workload family of `circuits.jl:61-67`).

    Single-qubit gates come from numpy's PCG64 stream seeded with `seed`; per
    cycle and qubit in qubit order one draw picks among the allowed gates."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rows * cols
    c = Circuit(n)
    pats = grid_edge_patterns(rows, cols)
    choices = ["sx", "sy", "sw"]
    prev = [None] * (n + 1)
    for t in range(depth):
        for q in range(1, n + 1):
            allowed = [g for g in choices if g != prev[q]]
            g = allowed[int(rng.integers(len(allowed)))]
            prev[q] = g
            c.add(g, q)
        for (a, b) in pats[t % 4]:
            c.add("cz", a, b)
    if final_h:
        for q in range(1, n + 1):
            c.add("h", q)
    return c


def create_param_rqc_circuit(rows: int, cols: int, depth: int, seed: int | None = None) -> Circuit:
    """Parameterised grid circuit: per cycle Rz(a) Ry(b) Rz(c) on every qubit
    (angles uniform in [0, 2pi) from `seed`), then CZ on the staggered patterns."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rows * cols
    c = Circuit(n)
    pats = grid_edge_patterns(rows, cols)
    p = 0
    for t in range(depth):
        for q in range(1, n + 1):
            a, b, cc = rng.uniform(0.0, 2 * math.pi, size=3)
            c.add("rz", q, theta=float(a), param=p)
            c.add("ry", q, theta=float(b), param=p + 1)
            c.add("rz", q, theta=float(cc), param=p + 2)
            p += 3
        for (x, y) in pats[t % 4]:
            c.add("cz", x, y)
    return c
