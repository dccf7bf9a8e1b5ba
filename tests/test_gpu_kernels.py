"""GPU parity of each contraction kernel family on single-step programs.

Every case is one `ncon` whose legs are split into batch (H), contracted (K),
free-A (M) and free-B (N) labels, with the legs of A, B and C in random orders
(the permutation the kernels fold into their loads).  The result must equal the
oracle's TTGT ncon (itself pinned to np.einsum) within 1e-12 normwise for
complex128 and 5e-6 for complex64."""

import numpy as np
import pytest

import qxoracle as O
from qxhip import Executor

pytestmark = pytest.mark.gpu


def single_step(seed, nh, nm, nn, nk, dims=None, c_rows_first=False):
    rng = np.random.default_rng(seed)
    labs = list(range(1, nh + nm + nn + nk + 1))
    H, M, N, K = labs[:nh], labs[nh:nh + nm], labs[nh + nm:nh + nm + nn], labs[nh + nm + nn:]
    al = H + M + K
    bl = H + K + N
    cl = H + M + N
    rng.shuffle(al)
    rng.shuffle(bl)
    rng.shuffle(cl)
    if c_rows_first and (H or M):
        first = (H + M)[0]
        cl.remove(first)
        cl.insert(0, first)
    dims = dims or {}
    d = {l: dims.get(l, 2) for l in labs}
    A = rng.normal(size=[d[l] for l in al]) + 1j * rng.normal(size=[d[l] for l in al])
    B = rng.normal(size=[d[l] for l in bl]) + 1j * rng.normal(size=[d[l] for l in bl])
    fmt = lambda ls: ",".join(map(str, ls)) if ls else "0"
    dsl = "\n".join([
        "# version: 0.4.0",
        f"load A data_1 {fmt([d[l] for l in al])}" if al else "load A data_1",
        f"load B data_2 {fmt([d[l] for l in bl])}" if bl else "load B data_2",
        f"ncon C {fmt(cl)} A {fmt(al)} B {fmt(bl)}",
        "save output C",
    ]) + "\n"
    data = {"data_1": A, "data_2": B}
    ref = O.ncon(A, al, B, bl, cl).reshape(-1, order="F")
    return dsl, data, ref


def run(dsl, data, precision="c128"):
    ex = Executor(dsl, data, precision=precision)
    out = ex.amplitudes([""])
    kern = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]["kernel"]
    ex.close()
    return np.asarray(out).reshape(-1), kern


def err(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


GEMM_CASES = [(0, 12, 5, 5), (0, 5, 3, 3), (0, 7, 4, 4), (0, 8, 5, 5), (1, 6, 4, 5), (2, 6, 5, 3), (0, 6, 6, 6),
              (0, 9, 7, 4), (1, 5, 8, 6), (0, 12, 7, 5), (0, 7, 3, 8),
              (0, 6, 6, 12), (1, 7, 5, 10)]


@pytest.mark.parametrize("case", range(len(GEMM_CASES)))
@pytest.mark.parametrize("seed", [0, 1])
def test_gemm_kernel(case, seed):
    nh, nm, nn, nk = GEMM_CASES[case]
    dsl, data, ref = single_step(100 * case + seed, nh, nm, nn, nk)
    out, kern = run(dsl, data)
    assert kern == "gemm"
    assert err(out, ref) < 1e-12


@pytest.mark.parametrize("shape", [(0, 0, 0, 10), (3, 0, 0, 9), (0, 2, 2, 8), (1, 3, 0, 12), (6, 0, 0, 8)])
def test_reduce_kernel(shape):
    nh, nm, nn, nk = shape
    dsl, data, ref = single_step(7, nh, nm, nn, nk)
    out, kern = run(dsl, data)
    assert kern in ("reduce", "bits")      # lanes along K, or along C with K split over workgroups
    assert err(out, ref) < 1e-12


@pytest.mark.parametrize("shape", [(0, 10, 1, 1), (2, 8, 2, 2), (4, 4, 4, 0), (0, 11, 0, 2), (8, 0, 2, 1),
                                   (0, 0, 0, 0), (1, 0, 0, 0), (0, 3, 3, 0), (12, 0, 0, 0)])
def test_bits_kernel(shape):
    nh, nm, nn, nk = shape
    dsl, data, ref = single_step(11, nh, nm, nn, nk)
    out, kern = run(dsl, data)
    assert kern == "bits"
    assert err(out, ref) < 1e-12


@pytest.mark.parametrize("shape", [(0, 14, 1, 1), (0, 16, 2, 2), (2, 12, 3, 3), (1, 13, 0, 2), (0, 12, 3, 0),
                                   (3, 10, 2, 1), (0, 14, 3, 3), (1, 12, 3, 2), (0, 14, 5, 4),
                                   (1, 12, 4, 4), (0, 13, 6, 4), (2, 11, 5, 4), (3, 12, 2, 3)])
def test_rows_kernel(shape):
    nh, nm, nn, nk = shape
    dsl, data, ref = single_step(13 + sum(shape), nh, nm, nn, nk, c_rows_first=True)
    out, kern = run(dsl, data)
    assert kern == "rows"
    assert err(out, ref) < 1e-12


def test_radix_kernel_odd_dims():
    dsl, data, ref = single_step(5, 1, 2, 2, 2, dims={1: 3, 2: 5, 5: 3, 6: 7})
    out, kern = run(dsl, data)
    assert kern == "radix"
    assert err(out, ref) < 1e-12


@pytest.mark.parametrize("case", [0, 2, 5])
def test_gemm_kernel_c64(case):
    nh, nm, nn, nk = GEMM_CASES[case]
    dsl, data, ref = single_step(300 + case, nh, nm, nn, nk)
    out, kern = run(dsl, data, "c64_4m")
    assert err(out, ref) < 5e-6


MFMA_CASES = [(0, 6, 6, 6), (0, 12, 7, 5), (0, 6, 6, 12), (1, 7, 6, 6), (0, 9, 8, 4)]


@pytest.mark.parametrize("case", MFMA_CASES)
@pytest.mark.parametrize("precision,tol", [("c64_4m", 2e-5), ("c64_3m", 5e-5)])
def test_gemm_mfma_c64(case, precision, tol):
    """complex64 GEMM steps on FP32 matrix cores, four or three real products."""
    dsl, data, ref = single_step(500 + sum(case), *case)
    out, kern = run(dsl, data, precision)
    assert kern == "gemm"
    assert err(out, ref) < tol


@pytest.mark.parametrize("case", MFMA_CASES)
def test_gemm_mfma_c128(case):
    """complex128 GEMM steps on FP64 matrix cores (the FP64 result layout differs
    from FP32's)."""
    dsl, data, ref = single_step(900 + sum(case), *case)
    ex = Executor(dsl, data, kernel=5)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["variant"] == "k_contract_gemm_mfma"
    assert err(out, ref) < 1e-12


FORCE_SHAPES = [(0, 12, 5, 5), (1, 7, 5, 10), (0, 10, 2, 3), (2, 8, 2, 2), (0, 2, 2, 8), (0, 6, 6, 6)]


@pytest.mark.parametrize("choice,prec", [(c, "c128") for c in list(range(9)) + [11]] +
                         [(5, "c64_4m"), (5, "c64_3m"), (7, "c64_4m"), (7, "c64_3m"), (8, "c64_4m"),
                          (11, "c64_4m"), (11, "c64_3m")])
@pytest.mark.parametrize("shape", FORCE_SHAPES)
def test_forced_kernel_matches_oracle(shape, choice, prec):
    """Every kernel the autotuner may pick computes the same contraction: each
    configuration (0 gather, 1 rows, 2-4 FP tiles, 5/7 matrix cores, 6 reduce, 8 rocBLAS)
    forced onto the step; the gather kernel and the GEMM tiles run any power-of-two
    step (rocBLAS only plain matrix layouts, else the step keeps its kernel)."""
    dsl, data, ref = single_step(700 + sum(shape), *shape, c_rows_first=True)
    ex = Executor(dsl, data, precision=prec, kernel=choice)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    kern = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]["kernel"]
    ex.close()
    assert err(out, ref) < (5e-5 if prec != "c128" else 1e-12)
    if choice in (0, 2, 3, 4, 5, 7):
        assert kern == ("bits" if choice == 0 else "gemm")


def plain_step(seed, nh, nm, nn, nk, a_kfirst, b_nfirst, c_nfirst):
    """One ncon whose operands are plain (batched) column-major matrices: each class's
    legs consecutive and in order, the batch legs slowest."""
    labs = list(range(1, nh + nm + nn + nk + 1))
    H, M, N, K = labs[:nh], labs[nh:nh + nm], labs[nh + nm:nh + nm + nn], labs[nh + nm + nn:]
    al = (K + M if a_kfirst else M + K) + H
    bl = (N + K if b_nfirst else K + N) + H
    cl = (N + M if c_nfirst else M + N) + H
    rng = np.random.default_rng(seed)
    A = rng.normal(size=[2] * len(al)) + 1j * rng.normal(size=[2] * len(al))
    B = rng.normal(size=[2] * len(bl)) + 1j * rng.normal(size=[2] * len(bl))
    fmt = lambda ls: ",".join(map(str, ls))
    dsl = "\n".join([
        "# version: 0.4.0",
        f"load A data_1 {fmt([2] * len(al))}",
        f"load B data_2 {fmt([2] * len(bl))}",
        f"ncon C {fmt(cl)} A {fmt(al)} B {fmt(bl)}",
        "save output C",
    ]) + "\n"
    return dsl, {"data_1": A, "data_2": B}, O.ncon(A, al, B, bl, cl).reshape(-1, order="F")


PLAIN_SHAPES = [(0, 6, 5, 7), (1, 5, 6, 4), (2, 4, 4, 8), (0, 9, 4, 6)]


@pytest.mark.parametrize("layout", range(8))
@pytest.mark.parametrize("shape", PLAIN_SHAPES)
@pytest.mark.parametrize("prec", ["c128", "c64_4m"])
def test_rocblas_plain_layouts(shape, layout, prec):
    """The rocBLAS candidate on every transpose / operand-order combination of plain
    matrix layouts (A as M x K or K x M, B as K x N or N x K, C as M x N or N x M)."""
    dsl, data, ref = plain_step(900 + layout + 10 * sum(shape), *shape, layout & 1, layout >> 1 & 1, layout >> 2)
    ex = Executor(dsl, data, precision=prec, kernel=8)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["kernel"] == "blas"
    assert err(out, ref) < (5e-6 if prec != "c128" else 1e-12)


@pytest.mark.parametrize("shape,layout", [((0, 10, 10, 10), l) for l in range(8)] +
                         [((1, 10, 11, 10), 3), ((1, 10, 10, 11), 4)])
def test_plain_gemm_split_bf16_layouts(shape, layout):
    """The hand-written split bf16 product (M, N, K >= 1024) on every transpose /
    operand-order combination (1024^3: the K range split over two workgroups per tile),
    with a batch leg and an uneven tile count: normwise 1e-4 against the oracle (the
    dropped lo.lo term and the split residual are ~2^-16 of each product)."""
    dsl, data, ref = plain_step(970 + layout + 10 * sum(shape), *shape, layout & 1, layout >> 1 & 1, layout >> 2)
    ex = Executor(dsl, data, precision="c64_bf16x3")
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["variant"].startswith("k_plain_gemm_bf16x3"), st["variant"]
    assert err(out, ref) < 1e-4


@pytest.mark.parametrize("layout", range(8))
@pytest.mark.parametrize("shape", PLAIN_SHAPES + [(0, 4, 4, 12), (3, 4, 5, 4), (0, 5, 4, 13)])
@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64_4m", 5e-6), ("c64_3m", 2e-5)])
def test_plain_gemm_layouts(shape, layout, prec, tol):
    """The hand-written matrix-core GEMM (kernel 11) on every transpose / operand-order
    combination of plain matrix layouts, with partial C tiles (16 or 32 rows or columns
    against tiles of 64 or 128) and long K ranges split over workgroups (4096 and 8192
    values into 256 or 512 outputs); complex128 4M, complex64 4M and 3M."""
    dsl, data, ref = plain_step(950 + layout + 10 * sum(shape), *shape, layout & 1, layout >> 1 & 1, layout >> 2)
    ex = Executor(dsl, data, precision=prec, kernel=11)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["kernel"] == "plain_gemm", st["config"]
    assert err(out, ref) < tol


@pytest.mark.parametrize("layout", range(8))
@pytest.mark.parametrize("shape", PLAIN_SHAPES + [(0, 4, 4, 12), (0, 5, 4, 13)])
def test_plain_gemm_complex128_three_products(shape, layout):
    """The complex128 plain GEMM as three real products per complex product (kernel
    12) on every layout: normwise 1e-12 against the oracle, and per element within
    1e-12 of the largest element (the Gauss form's rounding scales with |a| |b|)."""
    dsl, data, ref = plain_step(960 + layout + 10 * sum(shape), *shape, layout & 1, layout >> 1 & 1, layout >> 2)
    ex = Executor(dsl, data, precision="c128", kernel=12)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["kernel"] == "plain_gemm" and ",3m" in st["variant"], st["variant"]
    assert err(out, ref) < 1e-12
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref))


def test_plain_gemm_split_k_in_graph_replays():
    """A plain GEMM whose K range is split over workgroups (partials summed by a second
    kernel) captured into the call's graph: every replay equals the eager result bit
    for bit, with another plan's eager calls on the device between the replays."""
    dsl, data, ref = plain_step(81, 0, 5, 4, 13, 0, 1, 0)
    eager = Executor(dsl, data, kernel=11)
    want = np.asarray(eager.amplitudes([""])).reshape(-1)
    st = next(b for b in eager.stats()["batches"] if b["active"])["steps"][0]
    assert ",ks" in st["variant"], st["variant"]
    graphed = Executor(dsl, data, kernel=11, graph=1)
    other_dsl, other_data, other_ref = plain_step(82, 1, 7, 6, 8, 1, 0, 1)
    other = Executor(other_dsl, other_data)
    for rep in range(3):
        got = np.asarray(graphed.amplitudes([""])).reshape(-1)
        assert np.array_equal(got, want), f"replay {rep}"
        assert err(np.asarray(other.amplitudes([""])).reshape(-1), other_ref) < 1e-12
    assert err(want, ref) < 1e-12
    for ex in (eager, graphed, other):
        ex.close()


def test_rocblas_step_in_graph_replays():
    """A rocBLAS step captured into the call's graph: every replay equals the oracle."""
    dsl, data, ref = plain_step(77, 1, 7, 6, 8, 1, 0, 1)
    ex = Executor(dsl, data, kernel=8, graph=1)
    for _ in range(3):
        out = np.asarray(ex.amplitudes([""])).reshape(-1)
        assert err(out, ref) < 1e-12
    ex.close()


def test_split_bf16_product_in_graph_replays():
    """The split bf16 product (library products of the hi / lo planes) captured into
    the call's graph: every replay, with eager calls of another plan on the device
    between the replays, equals the eager result bit for bit."""
    import qxtest
    dsl, data, ref = qxtest.interleaved_gemm_program(nk=10, nm=10, nn=10, seed=6)
    eager = Executor(dsl, data, precision="c64_bf16x3")
    want = np.asarray(eager.amplitudes([""])).reshape(-1)
    graphed = Executor(dsl, data, precision="c64_bf16x3", graph=1)
    other_dsl, other_data, other_ref = plain_step(78, 1, 7, 6, 8, 1, 0, 1)
    other = Executor(other_dsl, other_data, kernel=8)
    for rep in range(3):
        got = np.asarray(graphed.amplitudes([""])).reshape(-1)
        assert np.array_equal(got, want), f"replay {rep}"
        assert err(np.asarray(other.amplitudes([""])).reshape(-1), other_ref) < 1e-12
    assert err(want, ref) < 1e-4
    for ex in (eager, graphed, other):
        ex.close()


@pytest.mark.parametrize("view,prec,a_data", [(False, "c128", False), (True, "c128", False), (False, "c64_4m", False),
                                              (False, "c128", True)])
def test_regrouped_contraction_matches_oracle(view, prec, a_data):
    """A large contraction whose operands the lowering regroups into plain matrices
    (run by the hand-written plain GEMM), also around a data operand's fixed layout;
    with sliced operands the layouts stay as written."""
    import qxtest
    dsl, data, ref = qxtest.interleaved_gemm_program(view_a=view, a_data=a_data)
    ex = Executor(dsl, data, precision=prec)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][-1]
    ex.close()
    assert (st["kernel"] == "plain_gemm") == (not view)
    assert err(out, ref) < (5e-6 if prec != "c128" else 1e-12)


def test_split_bf16_product_matches_oracle():
    """c64_bf16x3 on a 1024^3 contraction: three bf16 products of the hi / lo split,
    accumulated in FP32, on the hand-written bf16 matrix-core kernel (and on the vendor
    BLAS as the tuned alternative).  The dropped lo.lo term and the split residual are ~2^-16 of
    each product, so the normwise error stays near 1e-5 (bar: 1e-4), against 1e-6-ish
    for the FP32 4M product of the same step."""
    import qxtest
    dsl, data, ref = qxtest.interleaved_gemm_program(nk=10, nm=10, nn=10, seed=5)
    ex = Executor(dsl, data, precision="c64_bf16x3")
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][-1]
    ex.close()
    assert st["kernel"] == "plain_gemm" and st["variant"].startswith("k_plain_gemm_bf16x3"), st["variant"]
    e_split = err(out, ref)
    # the same split product on the vendor BLAS (planes in the workspace, 12 library GEMMs)
    ex = Executor(dsl, data, precision="c64_bf16x3", kernel=8)
    lib = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][-1]
    ex.close()
    assert st["kernel"] == "blas"
    assert err(lib, ref) < 1e-4 and err(out, lib) < 2e-5
    assert e_split < 1e-4
    ex = Executor(dsl, data, precision="c64_4m")
    e_4m = err(np.asarray(ex.amplitudes([""])).reshape(-1), ref)
    ex.close()
    assert e_4m < e_split < 200 * max(e_4m, 1e-7)


SWAP_SHAPES = [(1, 2, 12, 2), (0, 3, 13, 1), (2, 1, 10, 3), (0, 2, 14, 4)]


@pytest.mark.parametrize("shape", SWAP_SHAPES)
@pytest.mark.parametrize("prec", ["c128", "c64_4m"])
def test_rows_kernel_over_swapped_operands(shape, prec):
    """A skinny A against a long B: the row kernel forced onto the step runs over B's
    free legs with the operands exchanged (C = B.A), and matches the oracle."""
    dsl, data, ref = single_step(300 + sum(shape), *shape)
    ex = Executor(dsl, data, precision=prec, kernel=1)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["kernel"] == "rows" and st["config"].endswith("/swap")
    assert err(out, ref) < (5e-6 if prec != "c128" else 1e-12)


SKINNY_SHAPES = [(1, 3, 1, 12), (0, 4, 0, 13), (2, 2, 1, 12), (1, 5, 0, 12), (0, 0, 1, 14)]


@pytest.mark.parametrize("shape", SKINNY_SHAPES)
@pytest.mark.parametrize("prec", ["c128", "c64_4m"])
def test_skinny_contraction_matches_oracle(shape, prec):
    """Long K into at most 32 outputs per batch element (kernel 9): every lane keeps all
    (m, n) sums of its K slice, split K partials summed after."""
    dsl, data, ref = single_step(500 + sum(shape), *shape)
    ex = Executor(dsl, data, precision=prec, kernel=9)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["variant"].startswith("k_contract_skinny")
    assert err(out, ref) < (5e-6 if prec != "c128" else 1e-12)


@pytest.mark.parametrize("prec,tol", [("c128", 1e-12), ("c64_4m", 5e-6), ("c64_bf16x3", 1e-4)])
def test_gathered_library_gemm_matches_oracle(prec, tol):
    """A 2^34-flop step with shuffled leg orders (kernel 10): A and B gathered into plain
    copies, the hand-written plain GEMM (the split bf16 product for c64_bf16x3), C
    gathered back."""
    dsl, data, ref = single_step(811, 0, 10, 10, 11)
    ex = Executor(dsl, data, precision=prec, kernel=10)
    out = np.asarray(ex.amplitudes([""])).reshape(-1)
    st = next(b for b in ex.stats()["batches"] if b["active"])["steps"][0]
    ex.close()
    assert st["kernel"] == "plain_gemm" and st["variant"].endswith("+gather"), st["variant"]
    assert ("bf16x3" in st["variant"]) == (prec == "c64_bf16x3")
    assert err(out, ref) < tol
