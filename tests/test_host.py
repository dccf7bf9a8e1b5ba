"""Host-side parity with the reference's own unit tests (no GPU)."""

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import (TensorCache, build_compute_graph, circuits, compute_graph, convert_to_tnc, planning,
                   simulation, workloads)


# ------------------------------------------------------------------ test/test_tn_conversion.jl:6-28
def test_tn_conversion_counts():
    circ = circuits.create_test_circuit()
    tnc = convert_to_tnc(circ, no_input=True, no_output=True, decompose=False)
    assert tnc.qubits() == 3 and len(tnc) == 3
    assert all(i != o for i, o in zip(tnc.input_indices, tnc.output_indices))
    tnc = convert_to_tnc(circ, no_input=True, no_output=True, decompose=True)
    assert len(tnc) == 5                        # each CX split in two by SVD
    tnc = convert_to_tnc(circ, no_input=False, no_output=True, decompose=False)
    assert len(tnc) == 6
    first = tnc.tensor_data(tnc.keys()[0])
    assert np.array_equal(first, circuits.gate_matrix(circ.ops[0]))


def test_hyperedges_of_decomposed_ghz3():
    """Decomposed GHZ-3 with inputs and outputs: the hyperedge census is
    {4 tensors: 2 edges, 2 tensors: 4 edges} (test/test_contraction_planning.jl:39-44)."""
    tnc = convert_to_tnc(circuits.create_test_circuit(), decompose=True)
    holders = tnc.tn.index_holders()
    census = {}
    for i, h in holders.items():
        census[len(h)] = census.get(len(h), 0) + 1
    assert census == {4: 2, 2: 4}
    assert len(planning.line_graph(tnc)) == 6


# ------------------------------------------------------------------ test/test_compute_graph.jl:5-24
def test_tensor_cache_dedupe_thresholds():
    tc = TensorCache()
    rng = np.random.default_rng(0)
    a = rng.random((3, 4, 5))
    sym = tc.push(a)
    assert sym == tc.push(a)
    assert sym != tc.push(rng.random((3, 4, 5)))
    assert sym == tc.push(a + 0.1 * np.finfo(float).eps)
    assert sym != tc.push(a + 1.1 * np.finfo(float).eps)
    assert np.array_equal(tc[sym], a)
    assert len(tc) == 3
    assert sym == "data_1"
    with pytest.raises(KeyError):
        tc["data_99"]


def test_compute_graph_node_count():
    """nodes = tensors + plan + 1 save (test/test_compute_graph.jl:26-34), plus one
    node per view when slicing."""
    circ = circuits.create_test_circuit()
    tnc = convert_to_tnc(circ)
    plan = planning.contraction_plan(tnc, trials=2)
    g = build_compute_graph(tnc, plan)
    assert len(g) == len(tnc) + len(plan) + 1
    groups = planning.expand_slice_bonds_to_hyperindices(tnc, planning.greedy_slicing(tnc, plan, 2))
    g2 = build_compute_graph(tnc, plan, groups)
    views = sum(1 for c in g2.commands if c.kind == "view")
    assert views >= 2 and len(g2) == len(tnc) + len(plan) + 1 + views


# ------------------------------------------------------------------ test/test_contraction_planning.jl
@pytest.mark.parametrize("decompose", [False, True])
def test_plan_lengths_and_intermediates(decompose):
    tnc = convert_to_tnc(circuits.create_test_circuit(), decompose=decompose)
    for plan in (planning.contraction_plan(tnc, trials=2),
                 planning.order_to_contraction_plan(planning.elimination_order(planning.line_graph(tnc), 0), tnc),
                 planning.sweep_contraction_plan(tnc)):
        assert len(plan) == len(tnc) - 1
        produced = [c for _, _, c in plan]
        assert all(c.startswith("I") for c in produced) and len(set(produced)) == len(produced)
        consumed = [x for a, b, _ in plan for x in (a, b)]
        assert len(set(consumed)) == len(consumed)          # each tensor used once


def test_forty_z_gates_hyperedge_plan():
    """One qubit, 40 Z gates: one big hyperedge; the plan still contracts all 42
    tensors (test/test_contraction_planning.jl:90-98)."""
    c = circuits.Circuit(1)
    for _ in range(40):
        c.add("z", 1)
    tnc = convert_to_tnc(c)
    assert len(tnc) == 42
    plan = planning.contraction_plan(tnc, trials=2)
    assert len(plan) == 41
    g = build_compute_graph(tnc, plan)
    amps = O.amplitudes(g.to_dsl(), g.data, ["0", "1"])
    assert np.allclose(amps, [1, 0])


def test_slicing_scheme_groups_and_metadata():
    tnc = convert_to_tnc(circuits.create_rqc_circuit(3, 3, 8, seed=1))
    groups, plan, meta = planning.contraction_scheme(tnc, 3, trials=4)
    assert len(groups) == 3 and all(len(g) == 1 for g in groups)
    widths = meta["Slicing"]["log2 largest intermediate after slicing consecutive edges"]
    assert widths == sorted(widths, reverse=True)
    assert len(plan) == len(tnc) - 1


@pytest.mark.parametrize("nslice,e", [(0, 16), (2, 16), (3, 8)])
def test_roofline_cost_matches_lowered_step_shapes(nslice, e):
    # the planner's roofline time from its own index sets against the one computed from
    # the written DSL's ncon steps (the oracle's per-step B M N K, sliced legs 1):
    # sum over steps of max(8 BMNK / P, e B (MK + KN + MN) / BW), times the slice count
    tnc = convert_to_tnc(circuits.create_rqc_circuit(3, 3, 8, seed=3))
    groups, plan, meta = planning.contraction_scheme(tnc, nslice, trials=4, slicer="roofline",
                                                     element_bytes=e)
    dsl = build_compute_graph(tnc, plan, groups, metadata=meta).to_dsl()
    P, BW = 78.6e12, 8.0e12
    shapes = O.step_shapes(dsl)
    assert len(shapes) == len(plan)
    t = sum(max(8.0 * B * M * N * K / P, e * B * (M * K + K * N + M * N) / BW) for B, M, N, K in shapes)
    nsl = len(list(O.slice_assignments(O.parse(dsl))))
    assert nsl == 2 ** nslice
    sliced = frozenset(g[0] for g in groups)
    got = planning.PlanCost(tnc, plan).roofline_seconds(sliced, e)
    assert got == pytest.approx(t * nsl, rel=1e-12)
    assert meta["Slicing"]["Roofline seconds, all slices"] == pytest.approx(got, rel=1e-5)


def test_roofline_slicer_picks_and_cap():
    tnc = convert_to_tnc(circuits.create_rqc_circuit(4, 4, 10, seed=5))
    plan = planning.contraction_plan(tnc, trials=4, seed=42)
    pc = planning.PlanCost(tnc, plan)
    # one pick: the narrowest largest intermediate, then the least roofline time, over
    # every candidate index by exhaustion
    cands = [i for i in tnc.tn.dims if tnc.tn.dims[i] > 1]
    best = min(cands, key=lambda i: (round(pc.cost(frozenset([i]))[1], 6), pc.roofline_seconds(frozenset([i]))))
    assert planning.roofline_slicing(tnc, plan, 1) == [best]
    # a cap the plan already meets: each pick adds the least roofline time
    w0 = pc.cost()[1]
    got = planning.roofline_slicing(tnc, plan, 1, max_log2_size=w0)
    assert got == [min(cands, key=lambda i: pc.roofline_seconds(frozenset([i])))]
    # slicing more never lowers the total time below the unsliced plan, and the
    # roofline picks cost no more time than the width picks at the same width
    s_w = planning.greedy_slicing(tnc, plan, 3)
    s_r = planning.roofline_slicing(tnc, plan, 3)
    assert pc.cost(frozenset(s_r))[1] <= pc.cost(frozenset(s_w))[1] + 1e-9
    assert pc.roofline_seconds(frozenset()) <= pc.roofline_seconds(frozenset(s_r)) <= \
        pc.roofline_seconds(frozenset(s_w)) * (1 + 1e-12)
    with pytest.raises(ValueError):
        planning.contraction_scheme(tnc, 1, trials=1, slicer="treewidth")


# ------------------------------------------------------------------ DSL
def test_dsl_round_trip_and_labels():
    tnc = convert_to_tnc(circuits.create_rqc_circuit(2, 3, 6, seed=2))
    groups, plan, meta = planning.contraction_scheme(tnc, 2, trials=2)
    g = build_compute_graph(tnc, plan, groups, metadata=meta)
    text = g.to_dsl()
    assert text.startswith("# version: 0.4.0\n")
    cmds = compute_graph.parse_dsl(text)
    assert [(c.kind, c.name, tuple(map(str, c.args))) for c in cmds] == \
        [(c.kind, c.name, tuple(map(str, c.args))) for c in g.commands]
    for c in cmds:
        if c.kind == "ncon":
            cl, a, al, b, bl = c.args
            # labels numbered by first appearance in A then B; C = survivors of A then new of B
            seen = []
            for x in al + bl:
                if x not in seen:
                    seen.append(x)
            assert seen == list(range(1, len(seen) + 1))
            assert cl == [x for x in al if x in cl] + [x for x in bl if x in cl and x not in al]


def test_outputs_and_views_in_dsl():
    tnc = convert_to_tnc(circuits.create_ghz_circuit(4))
    groups, plan, meta = planning.contraction_scheme(tnc, 1, trials=2)
    g = build_compute_graph(tnc, plan, groups)
    outs = [c for c in g.commands if c.kind == "output"]
    assert [c.args for c in outs] == [(q, 2) for q in range(1, 5)]
    views = [c for c in g.commands if c.kind == "view"]
    assert views and all(c.args[1] == "v1" and c.args[3] == 2 for c in views)


# ------------------------------------------------------------------ test/test_simulation.jl:3-14
def test_bitstring_generators():
    allb = list(simulation.amplitudes_all(5))
    assert len(allb) == 32 and all(len(b) == 5 for b in allb)
    assert allb[1] == "00001" and allb[-1] == "11111"       # most significant first
    u = simulation.amplitudes_uniform(5, None, 10)
    assert len(u) == 10 and all(set(b) <= {"0", "1"} for b in u)
    assert simulation.amplitudes_uniform(7, 3, 4) == simulation.amplitudes_uniform(7, 3, 4)


def test_output_params_dict():
    p = simulation.output_params_dict(5, 3, "List", bitstrings=["111", "101"])
    assert p["method"] == "List" and p["params"]["num_samples"] == 2
    p = simulation.output_params_dict(5, 20, "Rejection", seed=42)
    assert p["params"]["num_samples"] == 20 and p["params"]["M"] == 0.0001
    with pytest.raises(ValueError):
        simulation.output_params_dict(5, 1, "Bogus")


def test_generate_and_load_simulation_files(tmp_path):
    circ = circuits.create_rqc_circuit(2, 2, 4, seed=3)
    prefix = str(tmp_path / "rqc")
    simulation.generate_simulation_files(circ, prefix, 1, output_args=simulation.output_params_dict(4, 5, seed=1))
    dsl, data, params = simulation.load_simulation_files(prefix)
    assert params["output"]["params"]["num_samples"] == 5
    bits = params["output"]["params"]["bitstrings"]
    amps = O.amplitudes(dsl, data, bits)
    sv = qxtest.statevector_amps(circ, bits)
    assert np.allclose(amps, sv, atol=1e-14)


def test_committed_workloads_match_their_generator():
    """The saved plans under workloads/ are what their generator writes."""
    for name in ("rqc_3x3_d16", "rqc_5x5_d20"):
        w = workloads.WORKLOADS[name]
        dsl, data, bits = w.load()
        assert bits == w.bitstrings()
        tnc = convert_to_tnc(w.circuit())
        plan = planning.sweep_contraction_plan(tnc, **w.plan_kwargs)
        body = lambda t: [ln for ln in t.splitlines() if ln and not ln.startswith("#")]
        assert body(build_compute_graph(tnc, plan).to_dsl()) == body(dsl)


def test_config1_workload_oracle_vs_statevector():
    w = workloads.WORKLOADS["rqc_3x3_d16"]
    dsl, data, bits = w.load()
    amps = O.amplitudes(dsl, data, bits[:16])
    sv = qxtest.statevector_amps(w.circuit(), bits[:16])
    assert np.max(np.abs(amps - sv) / np.abs(sv)) < 1e-12


def test_stage_blocks_keeps_qubit_order_and_limits_stages():
    circ = circuits.create_rqc_circuit(4, 4, 12, seed=3)
    tnc = convert_to_tnc(circ)
    blocks = planning.fuse_gates(tnc.gate_tensors, 4)
    staged = planning.stage_blocks(blocks, 8)
    assert sorted(map(str, staged)) == sorted(map(str, blocks))
    pos = {tuple(b[0]): i for i, b in enumerate(staged)}
    last = {}
    for names, qs in blocks:                     # every qubit keeps its block order
        for q in qs:
            if q in last:
                assert pos[last[q]] < pos[tuple(names)]
            last[q] = tuple(names)
    # greedy stages: a new stage starts when the next block does not fit
    touched = set()
    for names, qs in staged:
        if len(touched | set(qs)) > 8:
            touched = set()
        touched |= set(qs)
        assert len(touched) <= 8
