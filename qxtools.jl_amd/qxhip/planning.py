"""Contraction planning: plan triples and slice bond groups.

The reference obtains plans from FlowCutter / min-fill / netcon
(`src/contraction_planning.jl:90-176,318-367`), which are not available here; only
the *formats* it emits are on the contraction path:

* a plan is a list of `(A, B, C)` name triples, intermediates named `I<i>`
  (`src/contraction_planning.jl:386-448`, `:408`);
* slicing returns groups of indices, one group per sliced (hyper)edge
  (`expand_slice_bonds_to_hyperindices`, `:190-203`; `contraction_scheme`, `:219-299`).

This module produces both with its own planner: a randomised greedy pairwise
contraction search that is hyperedge-aware (an index is summed only when its last
two holders meet), best-of-N trials by total flops under a rank cap, followed by
greedy slicing that keeps the contraction order fixed and picks, one index at a
time, the index whose removal most reduces the largest intermediate (ties broken
by total sliced flops).
"""

from __future__ import annotations

import heapq
import math
from collections import OrderedDict

import numpy as np

from .tn import TensorNetwork, TensorNetworkCircuit


def _as_tn(tn) -> TensorNetwork:
    return tn.tn if isinstance(tn, TensorNetworkCircuit) else tn


def _open_of(tn, open_indices) -> tuple:
    """Open indices given explicitly, else those of a circuit's uncapped wires."""
    if open_indices:
        return tuple(open_indices)
    if isinstance(tn, TensorNetworkCircuit):
        return tuple(tn.open_indices())
    return tuple(getattr(tn, "open", ()) or ())


def _log2dim(d: int) -> float:
    return math.log2(d) if d > 0 else 0.0


def greedy_contraction_plan(tn, seed: int | None = None, alpha: float = 1.0,
                            temperature: float = 0.0, open_indices=()) -> list:
    """One randomised greedy pass.  Returns plan triples (A, B, C)."""
    open_indices = _open_of(tn, open_indices)
    tn = _as_tn(tn)
    rng = np.random.default_rng(seed)
    lg = {i: _log2dim(d) for i, d in tn.dims.items()}
    open_set = set(open_indices)
    cur = {n: frozenset(t.indices) for n, t in tn.tensors.items()}
    count: dict = {}
    holders: dict = {}
    for n, s in cur.items():
        for i in s:
            count[i] = count.get(i, 0) + 1
            holders.setdefault(i, set()).add(n)

    def size(s):
        return sum(lg[i] for i in s)

    def result(a, b):
        sa, sb = cur[a], cur[b]
        shared = sa & sb
        gone = {i for i in shared if count[i] == 2 and i not in open_set}
        return (sa | sb) - gone

    def score(a, b):
        sc = result(a, b)
        v = 2.0 ** size(sc) - alpha * (2.0 ** size(cur[a]) + 2.0 ** size(cur[b]))
        s = math.copysign(math.log1p(abs(v)), v)
        if temperature > 0:
            s -= temperature * (-math.log(-math.log(rng.uniform(1e-12, 1.0))))
        return s

    heap = []
    tick = 0

    def push_pairs(n):
        # sorted: the pairs' order (heap ties, noise draws) must not depend on the
        # interpreter's string hashing
        nonlocal tick
        seen = set()
        for i in sorted(cur[n]):
            for m in sorted(holders[i]):
                if m != n and m not in seen:
                    seen.add(m)
                    tick += 1
                    heapq.heappush(heap, (score(n, m), tick, n, m))

    for n in list(cur):
        push_pairs(n)

    plan = []
    step = 0
    while heap:
        _, _, a, b = heapq.heappop(heap)
        if a not in cur or b not in cur:
            continue
        sc = result(a, b)
        step += 1
        c = f"I{step}"
        for i in cur[a] | cur[b]:
            holders[i].discard(a)
            holders[i].discard(b)
        for i in cur[a] & cur[b]:
            count[i] -= 1
        for i in (cur[a] & cur[b]) - sc:
            count[i] = 0
        del cur[a]
        del cur[b]
        cur[c] = frozenset(sc)
        for i in sc:
            holders[i].add(c)
        plan.append((a, b, c))
        push_pairs(c)
    # disconnected components: join by outer products, smallest first
    rest = sorted(cur, key=lambda n: size(cur[n]))
    while len(rest) > 1:
        a, b = rest[0], rest[1]
        step += 1
        c = f"I{step}"
        sc = result(a, b)
        for i in cur[a] & cur[b]:
            count[i] -= 1
        del cur[a]
        del cur[b]
        cur[c] = frozenset(sc)
        plan.append((a, b, c))
        rest = sorted(cur, key=lambda n: size(cur[n]))
    return plan


def line_graph(tn) -> dict:
    """Adjacency of the network's line graph over (hyper)edges: two indices are
    adjacent when one tensor holds both (`convert_to_line_graph`,
    `src/contraction_planning.jl:47-63`, hyperedge variant)."""
    tn = _as_tn(tn)
    adj: dict = {i: set() for i in tn.dims}
    for t in tn.tensors.values():
        for i in t.indices:
            adj[i].update(j for j in t.indices if j != i)
    return adj


def elimination_order(adj: dict, seed: int | None = None, heuristic: str = "min_fill",
                      noise: float = 0.0) -> list:
    """Vertex elimination order by the min-fill or min-degree heuristic, ties and
    optional noise resolved with a seeded generator (the role of `qxg.min_fill`,
    `src/contraction_planning.jl:133`)."""
    rng = np.random.default_rng(seed)
    g = {v: set(n) for v, n in adj.items()}
    jitter = {v: float(rng.uniform()) for v in g}
    order = []

    def fill(v):
        nb = list(g[v])
        f = 0
        for x in range(len(nb)):
            gx = g[nb[x]]
            for y in range(x + 1, len(nb)):
                if nb[y] not in gx:
                    f += 1
        return f

    while g:
        best, best_key = None, None
        for v in g:
            d = len(g[v])
            primary = fill(v) if heuristic == "min_fill" else d
            key = (primary * (1.0 + noise * rng.uniform()) if noise else primary, d, jitter[v])
            if best_key is None or key < best_key:
                best, best_key = v, key
        nb = g.pop(best)
        for x in nb:
            g[x].discard(best)
            g[x].update(nb - {x})
        order.append(best)
    return order


def order_to_contraction_plan(order: list, tn, open_indices=()) -> list:
    """Edge elimination order -> plan triples (`order_to_contraction_plan`,
    `src/contraction_planning.jl:386-448`): eliminating an index contracts every
    current tensor holding it; more than two holders are merged smallest-pair-first
    (the reference uses netcon there, `:417-436`).  Intermediates are `I<i>`."""
    open_set = set(_open_of(tn, open_indices))
    tn = _as_tn(tn)
    lg = {i: _log2dim(d) for i, d in tn.dims.items()}
    cur = {n: set(t.indices) for n, t in tn.tensors.items()}
    count: dict = {}
    for s in cur.values():
        for i in s:
            count[i] = count.get(i, 0) + 1
    where: dict = {}
    for n, s in cur.items():
        for i in s:
            where.setdefault(i, set()).add(n)
    plan = []
    step = [0]

    def contract(a, b):
        sa, sb = cur.pop(a), cur.pop(b)
        shared = sa & sb
        gone = {i for i in shared if count[i] == 2 and i not in open_set}
        for i in shared:
            count[i] -= 1
        sc = (sa | sb) - gone
        step[0] += 1
        c = f"I{step[0]}"
        cur[c] = sc
        for i in sa | sb:
            where[i].discard(a)
            where[i].discard(b)
        for i in sc:
            where[i].add(c)
        plan.append((a, b, c))
        return c

    size = lambda s: sum(lg[i] for i in s)
    for idx in order:
        group = sorted(where.get(idx, ()))
        while len(group) > 1:
            best, best_key = None, None
            for x in range(len(group)):
                for y in range(x + 1, len(group)):
                    a, b = group[x], group[y]
                    sa, sb = cur[a], cur[b]
                    shared = sa & sb
                    sc = (sa | sb) - {i for i in shared if count[i] == 2 and i not in open_set}
                    key = (size(sc), -len(shared), a, b)
                    if best_key is None or key < best_key:
                        best, best_key = (a, b), key
            a, b = best
            c = contract(a, b)
            group = [g for g in group if g not in (a, b)] + [c]
    rest = sorted(cur, key=lambda n: size(cur[n]))
    while len(rest) > 1:
        contract(rest[0], rest[1])
        rest = sorted(cur, key=lambda n: size(cur[n]))
    return plan


class _PlanBuilder:
    """Emits plan triples while tracking index holders (hyperedge-aware)."""

    def __init__(self, tn, open_indices=()):
        open_indices = _open_of(tn, open_indices)
        tn = _as_tn(tn)
        self.lg = {i: _log2dim(d) for i, d in tn.dims.items()}
        self.open = set(open_indices)
        self.cur = {n: set(t.indices) for n, t in tn.tensors.items()}
        self.count: dict = {}
        for s in self.cur.values():
            for i in s:
                self.count[i] = self.count.get(i, 0) + 1
        self.plan: list = []
        self.n = 0

    def contract(self, a: str, b: str) -> str:
        sa, sb = self.cur.pop(a), self.cur.pop(b)
        shared = sa & sb
        gone = {i for i in shared if self.count[i] == 2 and i not in self.open}
        for i in shared:
            self.count[i] -= 1
        self.n += 1
        c = f"I{self.n}"
        self.cur[c] = (sa | sb) - gone
        self.plan.append((a, b, c))
        return c

    def chain(self, names: list) -> str | None:
        """Contract names left to right; returns the final name."""
        if not names:
            return None
        acc = names[0]
        for x in names[1:]:
            acc = self.contract(acc, x)
        return acc

    def size(self, name: str) -> float:
        return sum(self.lg[i] for i in self.cur[name])


def fuse_gates(gates: list, max_qubits: int) -> list:
    """Group a time-ordered gate list [(tensor names, qubits)] into blocks of at
    most `max_qubits` qubits, each qubit's gates staying in time order (a block
    closes when a gate would push it over the limit).  Returns [(names, qubits)]
    in an order that respects every qubit's time order."""
    active: dict = {}          # qubit -> block id
    blocks: dict = {}          # id -> [names, qubit set]
    emitted: list = []
    nid = 0

    def close(bid):
        names, qs = blocks.pop(bid)
        for q in qs:
            if active.get(q) == bid:
                del active[q]
        emitted.append((names, sorted(qs)))

    for names, qubits in gates:
        ids = {active[q] for q in qubits if q in active}
        union = set(qubits)
        for b in ids:
            union |= blocks[b][1]
        if len(union) > max_qubits:
            for b in sorted(ids):
                close(b)
            ids = set()
            union = set(qubits)
        nid += 1
        merged = []
        for b in sorted(ids):
            merged += blocks.pop(b)[0]
        blocks[nid] = [merged + list(names), union]
        for q in union:
            active[q] = nid
    for b in sorted(blocks):
        emitted.append((blocks[b][0], sorted(blocks[b][1])))
    return emitted


def stage_blocks(blocks: list, stage_qubits: int) -> list:
    """Reorder gate blocks [(names, qubits)] into stages that each touch at most
    `stage_qubits` qubits, keeping every qubit's blocks in their order.  Greedy: a
    stage takes ready blocks that fit, most already-touched qubits first (then the
    earliest).  Consecutive blocks of a stage can then run as one pass over the
    state (the executor's chain passes touch at most that many legs per pass)."""
    n = len(blocks)
    last: dict = {}
    preds = []
    for i, (_, qs) in enumerate(blocks):
        preds.append({last[q] for q in qs if q in last})
        for q in qs:
            last[q] = i
    done = [False] * n
    left = set(range(n))
    order = []
    while left:
        touched: set = set()
        while True:
            ready = [i for i in left if all(done[p] for p in preds[i])]
            fit = [i for i in ready if len(touched | set(blocks[i][1])) <= stage_qubits]
            if not fit:
                if not touched and ready:          # a block wider than a stage runs alone
                    fit = [min(ready)]
                else:
                    break
            pick = max(fit, key=lambda i: (len(touched & set(blocks[i][1])), -i))
            order.append(pick)
            done[pick] = True
            left.discard(pick)
            touched |= set(blocks[pick][1])
    return [blocks[i] for i in order]


def sweep_contraction_plan(tnc, max_block_qubits: int = 5, output_groups: int = 2,
                           stage_qubits: int | None = 10) -> list:
    """Circuit-aware plan for a closed network: fuse gates into blocks of at most
    `max_block_qubits` qubits, contract each block, apply blocks to the growing
    state (merging state components they touch), then absorb the output tensors
    last in `output_groups` outer-product groups.  Bitstring-independent work
    (everything before the outputs) is thereby one shared subtree.  Blocks are
    applied in stages of at most `stage_qubits` qubits (None: time order)."""
    b = _PlanBuilder(tnc)
    comp_of: dict = {}        # qubit -> current state component name
    for q, name in enumerate(tnc.input_tensors, start=1):
        comp_of[q] = name
    blocks = fuse_gates(tnc.gate_tensors, max_block_qubits)
    if stage_qubits:
        blocks = stage_blocks(blocks, stage_qubits)
    for names, qubits in blocks:
        blk = b.chain(list(names))
        comps = []
        for q in qubits:
            c = comp_of.get(q)
            if c is not None and c not in comps:
                comps.append(c)
        comps.sort(key=b.size)
        acc = blk
        for c in comps:
            acc = b.contract(c, acc) if b.size(c) >= b.size(acc) else b.contract(acc, c)
        for q, c in list(comp_of.items()):
            if c in comps or q in qubits:
                comp_of[q] = acc
    states = []
    for c in comp_of.values():
        if c not in states:
            states.append(c)
    states.sort(key=b.size)
    state = b.chain(states[::-1]) if len(states) > 1 else states[0]
    outs = list(tnc.output_tensors)
    if outs:
        k = max(1, min(output_groups, len(outs)))
        per = -(-len(outs) // k)
        groups = [outs[i:i + per] for i in range(0, len(outs), per)]
        for g in groups:
            og = b.chain(g)
            state = b.contract(state, og)
    return b.plan


class PlanCost:
    """Per-step index sets of a plan, for fast cost evaluation under slicing."""

    def __init__(self, tn, plan: list, open_indices=()):
        open_indices = _open_of(tn, open_indices)
        tn = _as_tn(tn)
        self.lg = {i: _log2dim(d) for i, d in tn.dims.items()}
        open_set = set(open_indices)
        cur = {n: set(t.indices) for n, t in tn.tensors.items()}
        count: dict = {}
        for s in cur.values():
            for i in s:
                count[i] = count.get(i, 0) + 1
        self.steps = []   # (all indices of A u B, indices of C)
        self.operands = []   # (indices of A, indices of B)
        for a, b, c in plan:
            sa, sb = cur.pop(a), cur.pop(b)
            shared = sa & sb
            gone = {i for i in shared if count[i] == 2 and i not in open_set}
            for i in shared:
                count[i] -= 1
            sc = (sa | sb) - gone
            cur[c] = sc
            self.steps.append((frozenset(sa | sb), frozenset(sc)))
            self.operands.append((frozenset(sa), frozenset(sb)))

    def cost(self, sliced=frozenset()) -> tuple:
        """(log2 total complex flops of one slice incl. 8x, log2 largest intermediate)"""
        lg = self.lg
        tot = 0.0
        mx = 0.0
        for allidx, cidx in self.steps:
            f = sum(lg[i] for i in allidx if i not in sliced)
            tot += 2.0 ** f
            s = sum(lg[i] for i in cidx if i not in sliced)
            mx = max(mx, s)
        return math.log2(8.0 * tot) if tot > 0 else 0.0, mx

    def roofline_seconds(self, sliced=frozenset(), element_bytes: int = 16,
                         flops_peak: float = 78.6e12, hbm_bw: float = 8.0e12) -> float:
        """Roofline time of the whole contraction over all slices: per step of one
        slice max(8 B M N K / flops_peak, e (|A| + |B| + |C|) / hbm_bw), summed over the
        steps and multiplied by the number of slices (the product of the sliced
        dimensions)."""
        lg = self.lg

        def size(idx):
            return 2.0 ** sum(lg[i] for i in idx if i not in sliced)

        t = 0.0
        for (allidx, cidx), (aidx, bidx) in zip(self.steps, self.operands):
            flops = 8.0 * size(allidx)
            nbytes = element_bytes * (size(aidx) + size(bidx) + size(cidx))
            t += max(flops / flops_peak, nbytes / hbm_bw)
        return t * 2.0 ** sum(lg[i] for i in sliced)


def plan_depth(plan: list) -> int:
    """Levels of the plan's dependency tree: the longest chain of steps each reading the
    previous one's result."""
    lvl: dict = {}
    for a, b, c in plan:
        lvl[c] = 1 + max(lvl.get(a, 0), lvl.get(b, 0))
    return max(lvl.values(), default=0)


def contraction_plan(tn, trials: int = 32, seed: int = 42, max_log2_size: float | None = None,
                     open_indices=(), min_fill_trials: int = 0, level_seconds: float = 5e-6) -> list:
    """Best plan (rank cap optional) over `trials` randomised greedy passes and
    `min_fill_trials` noisy min-fill line-graph elimination orders, by modelled device
    time: the roofline time of its steps (`PlanCost.roofline_seconds`) plus
    `level_seconds` per level of its dependency tree (`plan_depth`: the executor runs a
    level's small steps in one launch, and a launch costs a few microseconds however
    little it does).  Large plans are ranked by their arithmetic, small ones by their
    depth; total flops break ties.  level_seconds 0: by total flops alone."""
    open_indices = _open_of(tn, open_indices)
    tn = _as_tn(tn)
    rng = np.random.default_rng(seed)
    best, best_key = None, None
    cands = []
    for t in range(trials):
        if t == 0:
            alpha, temp = 1.0, 0.0
        else:
            alpha = float(np.exp(rng.uniform(np.log(0.1), np.log(4.0))))
            temp = float(np.exp(rng.uniform(np.log(0.01), np.log(1.0))))
        cands.append(("g", alpha, temp, int(rng.integers(1 << 31))))
    for t in range(min_fill_trials):
        cands.append(("m", float(rng.uniform(0.0, 0.3)), 0.0, int(rng.integers(1 << 31))))
    adj = line_graph(tn) if min_fill_trials else None
    for kind, a1, a2, sd in cands:
        if kind == "g":
            plan = greedy_contraction_plan(tn, seed=sd, alpha=a1, temperature=a2, open_indices=open_indices)
        else:
            order = elimination_order(adj, seed=sd, heuristic="min_fill", noise=a1)
            plan = order_to_contraction_plan(order, tn, open_indices)
        pc = PlanCost(tn, plan, open_indices)
        flops, mx = pc.cost()
        over = 0.0 if max_log2_size is None else max(0.0, mx - max_log2_size)
        if level_seconds > 0:
            key = (over, pc.roofline_seconds() + level_seconds * plan_depth(plan), flops, mx)
        else:
            key = (over, flops, mx)
        if best_key is None or key < best_key:
            best, best_key = plan, key
    return best


def greedy_slicing(tn, plan: list, num: int, open_indices=()) -> list:
    """Pick `num` indices to slice with the plan order fixed (the role of
    `greedy_treewidth_deletion` at `src/contraction_planning.jl:254`)."""
    open_indices = _open_of(tn, open_indices)
    tn = _as_tn(tn)
    pc = PlanCost(tn, plan, open_indices)
    open_set = set(open_indices)
    candidates = sorted(i for i in tn.dims if i not in open_set and tn.dims[i] > 1)
    sliced: list = []
    for _ in range(num):
        best, best_key = None, None
        for i in candidates:
            if i in sliced:
                continue
            s = frozenset(sliced + [i])
            flops, mx = pc.cost(s)
            key = (round(mx, 6), flops + len(s))
            if best_key is None or key < best_key:
                best, best_key = i, key
        if best is None:
            break
        sliced.append(best)
    return sliced


def roofline_slicing(tn, plan: list, num: int, open_indices=(), element_bytes: int = 16,
                     flops_peak: float = 78.6e12, hbm_bw: float = 8.0e12,
                     max_log2_size: float | None = None) -> list:
    """Pick `num` indices to slice with the plan order fixed, scoring each choice by
    the roofline time of all slices (`PlanCost.roofline_seconds`) instead of flops.

    Without `max_log2_size` each pick first minimises the largest intermediate, like
    `greedy_slicing`, and breaks ties by roofline time; with it, each pick minimises
    the excess over that size first and then the roofline time, so once the plan fits
    the picks add the least device time.  Replaces the treewidth cost of
    `contraction_scheme` (`src/contraction_planning.jl:219-299`)."""
    open_indices = _open_of(tn, open_indices)
    tn = _as_tn(tn)
    pc = PlanCost(tn, plan, open_indices)
    open_set = set(open_indices)
    candidates = sorted(i for i in tn.dims if i not in open_set and tn.dims[i] > 1)
    sliced: list = []
    for _ in range(num):
        best, best_key = None, None
        for i in candidates:
            if i in sliced:
                continue
            s = frozenset(sliced + [i])
            mx = pc.cost(s)[1]
            width = mx if max_log2_size is None else max(0.0, mx - max_log2_size)
            key = (round(width, 6), pc.roofline_seconds(s, element_bytes, flops_peak, hbm_bw))
            if best_key is None or key < best_key:
                best, best_key = i, key
        if best is None:
            break
        sliced.append(best)
    return sliced


def expand_slice_bonds_to_hyperindices(tn, bonds: list) -> list:
    """Group sliced indices by hyperedge (`src/contraction_planning.jl:190-203`).
    An index here already names its whole hyperedge, so each group has one index."""
    groups: list = []
    for b in bonds:
        if not any(b in g for g in groups):
            groups.append([b])
    return groups


def contraction_scheme(tn, num: int, trials: int = 32, seed: int = 42,
                       max_log2_size: float | None = None, min_fill_trials: int = 0,
                       slicer: str = "width", element_bytes: int = 16,
                       max_log2_slice_size: float | None = None, level_seconds: float = 5e-6) -> tuple:
    """(bond_groups, plan, metadata) as `contraction_scheme` returns
    (`src/contraction_planning.jl:219-299`).

    slicer: "width" (`greedy_slicing`, flops as the tie-break) or "roofline"
    (`roofline_slicing` with `element_bytes` per element and an optional cap on the
    largest intermediate of one slice).  level_seconds: see `contraction_plan` (0 ranks
    the candidates by flops alone)."""
    if slicer not in ("width", "roofline"):
        raise ValueError(f"unknown slicer {slicer!r}")
    open_indices = _open_of(tn, ())
    tn = _as_tn(tn)
    plan = contraction_plan(tn, trials=trials, seed=seed, max_log2_size=max_log2_size,
                            open_indices=open_indices, min_fill_trials=min_fill_trials,
                            level_seconds=level_seconds)
    pc = PlanCost(tn, plan, open_indices)
    flops0, mx0 = pc.cost()
    if num <= 0:
        sliced = []
    elif slicer == "roofline":
        sliced = roofline_slicing(tn, plan, num, open_indices, element_bytes=element_bytes,
                                  max_log2_size=max_log2_slice_size)
    else:
        sliced = greedy_slicing(tn, plan, num, open_indices)
    widths = [pc.cost(frozenset(sliced[:k + 1]))[1] for k in range(len(sliced))]
    groups = expand_slice_bonds_to_hyperindices(tn, sliced)
    meta = OrderedDict()
    meta["Determination of contraction plan"] = OrderedDict([
        ("Method used", "randomised greedy (best of trials by roofline time plus launch levels)"),
        ("Trials", trials), ("Seed used", seed),
        ("log2 flops", round(flops0, 3)), ("log2 largest intermediate", round(mx0, 3)),
        ("Hypergraph used", True),
    ])
    meta["Slicing"] = OrderedDict([
        ("Method used", "greedy largest-intermediate reduction, order fixed" if slicer == "width"
         else "greedy roofline-time reduction, order fixed"),
        ("Edges sliced", num),
        ("log2 largest intermediate after slicing consecutive edges", [round(w, 3) for w in widths]),
        ("Roofline seconds, all slices", float(f"{pc.roofline_seconds(frozenset(sliced), element_bytes):.6g}")),
    ])
    return groups, plan, meta
