"""Compute graph (the executor's program) and its `.qx` DSL text form.

Mirrors `build_compute_graph` (`src/compute_graph/compute_graph.jl:15-98`),
`TensorCache` (`src/compute_graph/tensor_cache.jl:6-111`) and the DSL specified in
`docs/src/users_guide.md:37-170`:

    load  <name> <data key> <d1,d2,...>
    output <name> <qubit> <dim>
    view  <new> <old> <slice var> <1-based leg> <dim>
    ncon  <C> <c labels> <A> <a labels> <B> <b labels>     (label 0 = scalar)
    save  <label> <name>

Output leg order of an `ncon` is A's surviving legs in A's order, then B's
remaining legs in B's order, as in the DSL example (`users_guide.md:76-89`).
Labels are numbered by first appearance in A then B.
"""

from __future__ import annotations

from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from .tn import TensorNetworkCircuit

DSL_VERSION = "0.4.0"


class TensorCache:
    """Deduplicating store of initial tensors (`tensor_cache.jl:6-81`): tensors of
    equal shape whose entries differ by less than eps(Float64) share a key
    `data_<i>` (`:51-69`); data is kept flattened column-major (`:53`)."""

    def __init__(self, label: str = "data_"):
        self.tensors: "OrderedDict[tuple, list]" = OrderedDict()
        self.key_dim_map: "OrderedDict[str, tuple]" = OrderedDict()
        self.id_val = 1
        self.label = label

    def next_symbol(self) -> str:
        s = f"{self.label}{self.id_val}"
        self.id_val += 1
        return s

    def push(self, data: np.ndarray) -> str:
        data = np.asarray(data)
        dim = tuple(int(d) for d in data.shape)
        flat = data.reshape(-1, order="F").astype(np.complex128)
        eps = np.finfo(np.float64).eps
        if dim in self.tensors:
            for sym, stored in self.tensors[dim]:
                if stored.size == 0 or np.max(np.abs(flat - stored)) < eps:
                    return sym
        sym = self.next_symbol()
        self.tensors.setdefault(dim, []).append((sym, flat))
        self.key_dim_map[sym] = dim
        return sym

    def __getitem__(self, sym: str) -> np.ndarray:
        if sym not in self.key_dim_map:
            raise KeyError(f"No symbol {sym} in cache")
        dim = self.key_dim_map[sym]
        for s, flat in self.tensors[dim]:
            if s == sym:
                return flat.reshape(dim, order="F")
        raise KeyError(sym)

    def __len__(self) -> int:
        return len(self.key_dim_map)

    def to_dict(self) -> dict:
        return {k: self[k] for k in self.key_dim_map}


class _NoDedupe(TensorCache):
    """TensorCache that never merges equal tensors."""

    def push(self, data: np.ndarray) -> str:
        data = np.asarray(data)
        dim = tuple(int(d) for d in data.shape)
        sym = self.next_symbol()
        self.tensors.setdefault(dim, []).append((sym, data.reshape(-1, order="F").astype(np.complex128)))
        self.key_dim_map[sym] = dim
        return sym


@dataclass
class Command:
    kind: str                 # load | output | view | ncon | save
    name: str
    args: tuple = ()


@dataclass
class ComputeGraph:
    commands: list
    data: dict
    metadata: "OrderedDict" = field(default_factory=OrderedDict)
    num_qubits: int = 0

    def __len__(self) -> int:
        """Number of nodes (every command is a node), `test/test_compute_graph.jl:33`."""
        return len(self.commands)

    def to_dsl(self) -> str:
        lines = [f"# version: {DSL_VERSION}"]
        lines += _metadata_lines(self.metadata)
        for c in self.commands:
            if c.kind == "load":
                key, dims = c.args
                lines.append(f"load {c.name} {key} {','.join(map(str, dims))}")
            elif c.kind == "output":
                q, d = c.args
                lines.append(f"output {c.name} {q} {d}")
            elif c.kind == "view":
                old, var, pos, d = c.args
                lines.append(f"view {c.name} {old} {var} {pos} {d}")
            elif c.kind == "ncon":
                cl, a, al, b, bl = c.args
                lines.append(f"ncon {c.name} {_lab(cl)} {a} {_lab(al)} {b} {_lab(bl)}")
            elif c.kind == "save":
                (src,) = c.args
                lines.append(f"save {c.name} {src}")
        return "\n".join(lines) + "\n"


def _lab(labels) -> str:
    return ",".join(map(str, labels)) if len(labels) else "0"


def _metadata_lines(meta, indent: int = 0) -> list:
    out = []
    for k, v in meta.items():
        pad = "  " * indent
        if isinstance(v, dict):
            out.append(f"# {pad}{k}:")
            out += _metadata_lines(v, indent + 1)
        elif isinstance(v, (list, tuple)):
            out.append(f"# {pad}{k}:")
            out += [f"# {pad}  - {x}" for x in v]
        else:
            out.append(f"# {pad}{k}: {v}")
    return out


def build_compute_graph(tnc: TensorNetworkCircuit, plan: list, bond_groups: list | None = None,
                        metadata=None, dedupe: bool = True) -> ComputeGraph:
    """Lower (network, plan, bond groups) to a compute graph
    (`src/compute_graph/compute_graph.jl:15-98`).  With `dedupe=False` every tensor
    gets its own data key (needed when gradients are taken per gate)."""
    tn = tnc.tn
    if hasattr(tnc, "open_indices") and callable(tnc.open_indices):
        open_set = set(tnc.open_indices())
    else:
        open_set = set(getattr(tnc, "open", ()) or ())
    legs = {n: list(t.indices) for n, t in tn.tensors.items()}
    count: dict = {}
    for n, l in legs.items():
        for i in l:
            count[i] = count.get(i, 0) + 1
    tc = TensorCache() if dedupe else _NoDedupe()
    node: "OrderedDict[str, Command]" = OrderedDict()
    for n, t in tn.tensors.items():                                   # :25-29
        node[n] = Command("load", n, (tc.push(t.data), tuple(t.data.shape)))
    oq = getattr(tnc, "output_qubits", None) or list(range(1, len(tnc.output_tensors) + 1))
    for q, o in zip(oq, tnc.output_tensors):                               # :32-35
        node[o] = Command("output", o, (q, int(tn[o].data.shape[0])))
    views = []
    renamed: dict = {}
    for gi, bg in enumerate(bond_groups or [], start=1):               # :39-58
        var = f"v{gi}"
        dim = tn.dims[bg[0]]
        related = [n for n in tn.tensors if any(b in legs[n] for b in bg)]
        for t in related:
            old = renamed.get(t, t)
            new = f"{old}_s"
            pos = next(j for j, i in enumerate(legs[t]) if i in bg) + 1
            views.append(Command("view", new, (old, var, pos, dim)))
            renamed[t] = new
    ncons = []
    cur = dict(legs)
    parent = set()

    def contract(a, b, c):
        la, lb = cur.pop(a), cur.pop(b)
        shared = set(la) & set(lb)
        gone = {i for i in shared if count[i] == 2 and i not in open_set}
        for i in shared:
            count[i] -= 1
        lc = [i for i in la if i not in gone] + [i for i in lb if i not in gone and i not in la]
        num: dict = {}
        for i in la + lb:
            num.setdefault(i, len(num) + 1)
        ncons.append(Command("ncon", c, ([num[i] for i in lc], renamed.get(a, a),
                                        [num[i] for i in la], renamed.get(b, b),
                                        [num[i] for i in lb])))
        cur[c] = lc
        parent.update((a, b))

    for a, b, c in plan:                                                # :61-74
        contract(a, b, c)
    roots = [n for n in cur]                                            # :77-90
    k = 0
    while len(roots) > 1:
        k += 1
        a, b = roots[0], roots[1]
        c = f"J{k}"
        contract(a, b, c)
        roots = [n for n in cur]
    root = roots[0]
    cmds = list(node.values()) + views + ncons + [Command("save", "output", (renamed.get(root, root),))]
    return ComputeGraph(cmds, tc.to_dict(), OrderedDict(metadata or {}), tnc.num_qubits)


def parse_dsl(text: str) -> list:
    """Parse `.qx` text into Commands (inverse of `ComputeGraph.to_dsl`)."""
    cmds = []
    for ln, raw in enumerate(text.splitlines(), start=1):
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        f = line.split()
        op = f[0]
        try:
            if op == "load":
                dims = tuple(int(x) for x in f[3].split(",")) if len(f) > 3 else ()
                cmds.append(Command("load", f[1], (f[2], dims)))
            elif op == "output":
                cmds.append(Command("output", f[1], (int(f[2]), int(f[3]))))
            elif op == "view":
                cmds.append(Command("view", f[1], (f[2], f[3], int(f[4]), int(f[5]))))
            elif op == "ncon":
                lab = lambda s: [] if s == "0" else [int(x) for x in s.split(",")]
                cmds.append(Command("ncon", f[1], (lab(f[2]), f[3], lab(f[4]), f[5], lab(f[6]))))
            elif op == "save":
                cmds.append(Command("save", f[1], (f[2],)))
            else:
                raise ValueError(f"unknown instruction {op!r}")
        except (IndexError, ValueError) as e:
            raise ValueError(f"DSL line {ln}: {raw!r}: {e}") from None
    return cmds
