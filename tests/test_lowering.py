"""Host-side checks of libqxhip (no GPU): symbols, DSL validation and error
statuses, kernel choice, and the arena reuse schedule (no two live values overlap)."""

import ctypes
import re

import pytest

import test_gpu_kernels as K
from qxhip import _native, workloads

HEADER = __import__("os").path.join(__import__("os").path.dirname(__file__), "..", "include", "qxhip.h")


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(qx_[a-z_]+)\s*\(", text))
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_native.SYMBOLS) <= declared


def _status(dsl):
    try:
        _native.inspect(dsl)
        return 0
    except _native.QxError as e:
        return e.status


BAD = {
    "unknown op": ("load A data_1 2\nfoo A\nsave output A\n", -2),
    "no save": ("load A data_1 2\n", -2),
    "bad dim": ("load A data_1 2,x\nsave output A\n", -2),
    "undefined input": ("load A data_1 2\nncon C 1 A 1 B 1\nsave output C\n", -2),
    "redefined": ("load A data_1 2\nload A data_2 2\nsave output A\n", -2),
    "repeated label": ("load A data_1 2,2\nload B data_2 2\nncon C 1 A 1,1 B 1\nsave output C\n", -3),
    "sum-out of A alone": ("load A data_1 2,2\nload B data_2 2\nncon C 1 A 1,2 B 1\nsave output C\n", -3),
    "label in neither": ("load A data_1 2\nload B data_2 2\nncon C 1,3 A 1 B 1\nsave output C\n", -3),
    "dim mismatch": ("load A data_1 2\nload B data_2 3\nncon C 0 A 1 B 1\nsave output C\n", -3),
    "rank mismatch": ("load A data_1 2,2\nload B data_2 2\nncon C 1 A 1 B 1\nsave output C\n", -3),
    "view dim": ("load A data_1 2\nview A_s A v1 1 3\nsave output A_s\n", -2),
    "view leg": ("load A data_1 2\nview A_s A v1 2 2\nsave output A_s\n", -2),
    "output dim": ("output o 1 3\nsave output o\n", -2),
    "key sizes": ("load A data_1 2\nload B data_1 2,2\nsave output A\n", -2),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_malformed_programs_return_status(case):
    dsl, code = BAD[case]
    assert _status(dsl) == code


def test_dsl_example_from_users_guide_lowers():
    """The two-qubit GHZ example of docs/src/users_guide.md:71-90 (views, hyper
    labels, a scalar result) is accepted."""
    dsl = """# version: 0.4.0
load t5 data_1 2
view t5_s t5 v2 1 2
load t1 data_2 2,2
view t1_s t1 v1 1 2
load t4 data_1 2
ncon I2 1 t1_s 1,2 t4 2
ncon t8 1,2 t5_s 1 I2 2
load t3 data_4 2
view t3_s t3 v1 1 2
ncon t9 1,2 t8 1,2 t3_s 2
output t6 1 2
view t6_s t6 v1 1 2
ncon t10 1,2 t9 1,2 t6_s 2
load t2 data_3 2,2,2
view t2_s t2 v1 3 2
view t2_s_s t2_s v2 2 2
output t7 2 2
ncon I1 2,3 t2_s_s 1,2,3 t7 1
ncon t11 0 t10 1,2 I1 1,2
save output t11
"""
    info = _native.inspect(dsl, batch=4)
    names = [s["name"] for s in info["steps"]]
    # t8 (outer product of t5_s and I2) and t9 (t8 times t3_s on a shared leg) form
    # one outer-product run
    assert names == ["I2", "t9", "t10", "I1", "t11"]
    t9 = info["steps"][1]
    assert t9["kernel"] == "outer" and t9["fused_steps"] == 2
    assert info["result_elems"] == 1


def _check_schedule(info):
    slots = info["slots"]
    for i, a in enumerate(slots):
        assert a["offset"] % 16 == 0
        assert a["offset"] + a["elems"] <= info["arena_bytes"] // (16 if info.get("c128", True) else 8)
        for b in slots[i + 1:]:
            live = a["start"] < b["end"] and b["start"] < a["end"]
            disjoint = a["offset"] + a["elems"] <= b["offset"] or b["offset"] + b["elems"] <= a["offset"]
            assert not live or disjoint, (a, b)


@pytest.mark.parametrize("name", ["rqc_3x3_d16", "rqc_5x5_d20"])
@pytest.mark.parametrize("batch", [1, 8, 256])
def test_arena_schedule_has_no_overlap(name, batch):
    dsl, data, bits = workloads.WORKLOADS[name].load()
    info = _native.inspect(dsl, batch=batch)
    _check_schedule(info)
    peak_live = 0
    ends = sorted({s["start"] for s in info["slots"]})
    for t in ends:
        peak_live = max(peak_live, sum(s["elems"] for s in info["slots"] if s["start"] <= t < s["end"]))
    assert info["arena_bytes"] >= 16 * peak_live
    # reuse must make the arena far smaller than the sum of all intermediates
    assert info["arena_bytes"] < 16 * sum(s["elems"] for s in info["slots"])


def test_sections_split_shared_and_bitstring_work():
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    info = _native.inspect(dsl, batch=256)
    secs = [s["section"] for s in info["steps"]]
    assert secs.count(2) < 40 < secs.count(0)   # outputs absorbed last by the sweep plan
    big = [s for s in info["steps"] if s["section"] == 2 and s["K"] >= 4096]
    # the state is regrouped into a plain K x M matrix, so the hand-written plain GEMM runs it
    assert big and big[0]["N"] == 256 and big[0]["kernel"] == "plain_gemm"


def test_kernel_choice_matches_gpu_test_shapes():
    for case in K.GEMM_CASES:
        dsl, _, _ = K.single_step(1, *case)
        assert _native.inspect(dsl)["steps"][0]["kernel"] == "gemm", case
    rows = [(0, 14, 1, 1), (0, 16, 2, 2), (2, 12, 3, 3), (1, 13, 0, 2), (0, 12, 3, 0), (3, 10, 2, 1),
            (0, 14, 3, 3), (1, 12, 3, 2), (0, 14, 5, 4), (1, 12, 4, 4), (0, 13, 6, 4), (2, 11, 5, 4),
            (3, 12, 2, 3)]
    for shape in rows:
        dsl, _, _ = K.single_step(13 + sum(shape), *shape, c_rows_first=True)
        assert _native.inspect(dsl)["steps"][0]["kernel"] == "rows", shape
    for shape in [(0, 0, 0, 10), (3, 0, 0, 9), (0, 2, 2, 8)]:
        dsl, _, _ = K.single_step(7, *shape)
        assert _native.inspect(dsl)["steps"][0]["kernel"] == "reduce", shape
    for shape in [(1, 3, 0, 12), (6, 0, 0, 8)]:        # C's bits have the unit stride: lanes along C
        dsl, _, _ = K.single_step(7, *shape)
        assert _native.inspect(dsl)["steps"][0]["kernel"] == "bits", shape
    dsl, _, _ = K.single_step(5, 1, 2, 2, 2, dims={1: 3, 2: 5, 5: 3, 6: 7})
    assert _native.inspect(dsl)["steps"][0]["kernel"] == "radix"


def test_complex64_wide_gemm_uses_matrix_cores():
    for case in K.MFMA_CASES:
        dsl, _, _ = K.single_step(1, *case)
        steps = _native.inspect(dsl, precision="c64_4m")["steps"]
        assert steps[0]["variant"] == "k_contract_gemm_mfma", case
        assert _native.inspect(dsl)["steps"][0]["variant"] != "k_contract_gemm_mfma"


@pytest.mark.parametrize("name", ["rqc_3x3_d16", "rqc_5x5_d20", "rqc_6x6_d12_s6"])
def test_autotune_candidates(name):
    """Every step lists its alternatives for the autotuner: the chosen kernel first,
    no duplicates, every GEMM-shaped step all three FP64 tiles."""
    w = workloads.WORKLOADS[name]
    dsl, data, bits = w.load()
    prec = "c128" if w.precision == "c128" else w.precision
    info = _native.inspect(dsl, precision=prec, batch=64)
    for st in info["steps"]:
        c = st["candidates"]
        assert c[0] == st["config"] and len(set(c)) == len(c)
        if st["kernel"] == "gemm":
            assert {"k_contract_gemm<128,16,16>", "k_contract_gemm<128,32,16>", "k_contract_gemm<64,64,16>",
                    "k_contract_gemm_mfma"} <= set(c)


def test_chains_fold_gate_block_runs():
    """The sweep plan of the 5x5 depth-20 workload applies gate blocks to a 2^25
    state; runs of them fold into chain passes (fewer passes over the state)."""
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    info = _native.inspect(dsl, batch=256)
    chains = [s for s in info["steps"] if s["kernel"] == "chain"]
    assert len(chains) >= 10 and all(s["fused_steps"] >= 2 for s in chains)
    big = [s for s in info["steps"] if s["bytes"] > 1e8]
    assert len(big) <= 20                      # 72 passes without chains or stages
    for s in chains:   # launch configurations of the one chain kernel
        assert all(c.startswith("k_contract_chain/") for c in s["candidates"]) and len(s["candidates"]) >= 3


def test_outer_product_runs_fold_into_one_step():
    """The output vectors of the 5x5 depth-20 plan (13 and 12 qubits' bitstring
    vectors, one outer product per qubit) are built by two outer-product runs, each
    one launch writing only the final 256 x 2^13 / 2^12 product; the factors it reads
    are the per-qubit vectors and the run's first operand."""
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    info = _native.inspect(dsl, batch=256)
    runs = [s for s in info["steps"] if s["kernel"] == "outer"]
    assert sorted(s["M"] for s in runs) == [256 << 12, 256 << 13]
    for s in runs:
        assert s["section"] == 2 and s["fused_steps"] >= 10 and len(s["reads"]) == s["fused_steps"] + 1
        assert s["candidates"] == ["k_outer_run"]
    assert not any(s["kernel"] == "rows" and s["section"] == 2 for s in info["steps"])
    _check_schedule(info)


def test_batched_small_levels_do_not_share_memory():
    """Small slice-free steps of one dependency level run in one launch, so within a
    level no step may write memory another step of the level reads or writes."""
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    info = _native.inspect(dsl, batch=256)
    slot = {s["name"]: (s["offset"], s["offset"] + s["elems"]) for s in info["slots"]}
    levels = {}
    for st in info["steps"]:
        if st["level"] >= 0:
            levels.setdefault(st["level"], []).append(st)
    assert len(levels) >= 5 and sum(len(v) for v in levels.values()) >= 300
    overlap = lambda a, b: a[0] < b[1] and b[0] < a[1]
    for lv, sts in levels.items():
        outs = [slot[s["name"]] for s in sts]
        reads = [slot[r] for s in sts for r in s["reads"] if r]
        for i, o in enumerate(outs):
            assert not any(overlap(o, x) for j, x in enumerate(outs) if j != i)
            assert not any(overlap(o, r) for r in reads)
        assert all(s["variant"] == "k_small_batch" for s in sts)


def test_split_bf16_precision_takes_large_plain_steps():
    """c64_bf16x3 runs plain steps with M, N, K >= 1024 as split bf16 products on the
    hand-written bf16 matrix-core GEMM; smaller steps keep the complex64 products, and
    no precision takes the vendor library by default (it stays a tuned candidate)."""
    import qxtest
    dsl, _, _ = qxtest.interleaved_gemm_program(nk=10, nm=10, nn=10)
    st = _native.inspect(dsl, precision="c64_bf16x3")["steps"][-1]
    assert st["kernel"] == "plain_gemm" and st["variant"].startswith("k_plain_gemm_bf16x3")
    assert any(c.startswith("rocblas") for c in st["candidates"])
    st = _native.inspect(dsl, precision="c64_4m")["steps"][-1]
    assert st["kernel"] == "plain_gemm" and not st["variant"].startswith("k_plain_gemm_bf16x3")
    dsl, _, _ = qxtest.interleaved_gemm_program(nk=11, nm=8, nn=8)   # M = N = 256
    st = _native.inspect(dsl, precision="c64_bf16x3")["steps"][-1]
    assert st["kernel"] != "blas" and "bf16x3" not in st["variant"]


def test_plain_matrix_steps_offer_rocblas():
    """Steps whose operands are plain (batched) column-major matrices list the rocBLAS
    GEMM among their candidates; shuffled leg orders do not."""
    for layout in range(8):
        dsl, _, _ = K.plain_step(layout, 1, 5, 6, 4, layout & 1, layout >> 1 & 1, layout >> 2)
        c = _native.inspect(dsl)["steps"][0]["candidates"]
        assert any(x.startswith("rocblas_gemm_strided_batched") for x in c), layout
    dsl, _, _ = K.single_step(5, 1, 5, 6, 4)
    assert not any(x.startswith("rocblas") for x in _native.inspect(dsl)["steps"][0]["candidates"])


def test_large_contraction_operands_regrouped_into_plain_matrices():
    """The lowering regroups the interleaved legs of a large contraction's operands
    (each produced for that contraction alone) into plain matrices, so the hand-written
    plain GEMM runs it; a viewed operand keeps its layout, and the step its tile kernel."""
    import qxtest
    dsl, _, _ = qxtest.interleaved_gemm_program()
    last = _native.inspect(dsl)["steps"][-1]
    assert last["kernel"] == "plain_gemm", last["config"]
    # A's contracted legs form one dense run and its free legs another
    ka = sorted(l[1] for l in last["klegs"])
    assert ka == [2 ** i * ka[0] for i in range(len(ka))]
    ma = sorted(l[1] for l in last["legs"] if l[1] and not l[2])
    assert ma == [2 ** i * ma[0] for i in range(len(ma))]
    dsl, _, _ = qxtest.interleaved_gemm_program(view_a=True)
    last = _native.inspect(dsl)["steps"][-1]
    assert last["kernel"] not in ("blas", "plain_gemm")
    assert not any(x.startswith("rocblas") for x in last["candidates"])
    # too small for the library GEMM: left in DSL order
    dsl, _, _ = qxtest.interleaved_gemm_program(nk=6, nm=6, nn=6)
    assert _native.inspect(dsl)["steps"][-1]["kernel"] not in ("blas", "plain_gemm")
    # a data operand keeps its layout (its shuffled K order): the intermediate follows it
    dsl, _, _ = qxtest.interleaved_gemm_program(a_data=True)
    assert _native.inspect(dsl)["steps"][-1]["kernel"] == "plain_gemm"


def _lds_groups(swz, bits, nb):
    """Distinct LDS bank groups (low nb address bits) that lanes differing in `bits`
    reach through the swizzle table (brute force over the 2^len(bits) lanes)."""
    seen = set()
    for m in range(1 << len(bits)):
        a = 0
        for i, b in enumerate(bits):
            if (m >> i) & 1:
                a ^= swz[b]
        seen.add(a & ((1 << nb) - 1))
    return len(seen)


def test_chain_lds_swizzle_is_a_bijection_that_spreads_banks():
    """Every chain tile's swizzle maps the 2^13 tile elements onto distinct LDS slots
    (brute force), keeps the low four tile bits plain, and lets the lanes of each
    step's first access cycle
    (16 lanes of 16-byte elements, differing in the step's lowest four free tile bits
    outside the wave bits of a wave-local chain, in the vector and the matrix-core
    steps alike) reach no fewer bank groups than the plain layout, and all 16 in the
    5x5 plan with its merged gate blocks (the plain layout: 3.5 on average; its chains
    are all wave-local, so the wave bits take three of the low free bits)."""
    dsl, data, bits = workloads.WORKLOADS["rqc_5x5_d20"].load()
    info = _native.inspect(dsl, batch=256)
    chains = [s for s in info["steps"] if s["kernel"] == "chain"]
    assert chains and sum(c["chain_tile"]["swizzled"] for c in chains) >= len(chains) // 2
    nb = 4
    groups, plain_groups = [], []
    for c in chains:
        t = c["chain_tile"]
        swz, pout = t["swz"], t["pout"]
        addr = set()
        for i in range(1 << 13):
            a = 0
            for b in range(13):
                if (i >> b) & 1:
                    a ^= swz[b]
            addr.add(a)
        assert addr == set(range(1 << 13))
        for b in range(nb):
            assert swz[b] == 1 << b
        plain = [1 << b for b in range(13)]
        used = swz if t["swizzled"] else plain
        wave = t["wbits"] if t["wave_local"] else []
        for q, h in zip(t["q"], t["hq"]):
            assert not set(wave) & (set(q) | set(h))
            free = [b for b in range(13) if b not in q and b not in h and b not in wave][:nb]
            got, ref = _lds_groups(used, free, nb), _lds_groups(plain, free, nb)
            assert got >= ref
            groups.append(got)
            plain_groups.append(ref)
    assert all(c["chain_tile"]["wave_local"] for c in chains)
    assert min(groups) == 1 << nb and sum(groups) > 4 * sum(plain_groups)
