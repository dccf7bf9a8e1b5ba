"""Drive the sanitizer build of libqxhip (simulated device) through a call sequence.

    LD_PRELOAD=<libasan.so> ASAN_OPTIONS=detect_leaks=0 \\
    QXHIP_LIB=qxtools.jl_amd/_build/libqxhip_san.so \\
    python qxtools.jl_amd/sanitize/run_sequence.py WORKLOAD [--fuse-slices 0] [--autotune 1]
        [--graph 1] [--calls 2] [--bits N] [--vjp] [--noalloc] [--precision c64_4m]

WORKLOAD is a saved plan of qxhip.workloads, or "sliced" (a small plan with two
sliced bonds).  Each call prints a line with the launches checked so far; any
out-of-reach operand aborts with "SIM DEVICE VIOLATION", any host memory error with
the sanitizer's report."""

from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from qxhip import Executor, _native, circuits, compute_graph, planning, simulation, tn, workloads  # noqa: E402


def sliced_plan():
    circ = circuits.create_rqc_circuit(2, 3, 6, seed=21)
    t = tn.convert_to_tnc(circ)
    plan = planning.contraction_plan(t, trials=8, seed=42)
    groups = planning.expand_slice_bonds_to_hyperindices(t, planning.greedy_slicing(t, plan, 2))
    g = compute_graph.build_compute_graph(t, plan, groups)
    return g.to_dsl(), g.data, simulation.amplitudes_uniform(6, 99, 8)


def plain_program(nh, nm, nn, nk, layout, seed=0):
    """One ncon of plain (batched) column-major matrices in transpose / order variant
    `layout` (bit 0: A stored K x M, bit 1: B stored N x K, bit 2: C stored N x M)."""
    import numpy as np

    labs = list(range(1, nh + nm + nn + nk + 1))
    H, M, N, K = labs[:nh], labs[nh:nh + nm], labs[nh + nm:nh + nm + nn], labs[nh + nm + nn:]
    al = (K + M if layout & 1 else M + K) + H
    bl = (N + K if layout & 2 else K + N) + H
    cl = (N + M if layout & 4 else M + N) + H
    rng = np.random.default_rng(seed)
    fmt = lambda ls: ",".join(map(str, ls))
    dsl = "\n".join(["# version: 0.4.0", f"load A data_1 {fmt([2] * len(al))}", f"load B data_2 {fmt([2] * len(bl))}",
                     f"ncon C {fmt(cl)} A {fmt(al)} B {fmt(bl)}", "save output C"]) + "\n"
    data = {"data_1": rng.normal(size=[2] * len(al)) + 0j, "data_2": rng.normal(size=[2] * len(bl)) + 0j}
    return dsl, data, [""]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("--fuse-slices", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=0)
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--calls", type=int, default=2)
    ap.add_argument("--bits", type=int, default=0, help="bitstrings per call (default: the workload's)")
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--vjp", action="store_true")
    ap.add_argument("--noalloc", action="store_true")
    ap.add_argument("--kernel", type=int, default=None, help="force a kernel choice (plan option)")
    a = ap.parse_args()
    if a.workload.startswith("plain:"):   # plain:nh,nm,nn,nk,layout
        nh, nm, nn, nk, layout = (int(x) for x in a.workload[6:].split(","))
        dsl, data, bits = plain_program(nh, nm, nn, nk, layout)
        prec = a.precision or "c128"
    elif a.workload == "sliced":
        dsl, data, bits = sliced_plan()
        prec = a.precision or "c128"
    else:
        w = workloads.WORKLOADS[a.workload]
        dsl, data, bits = w.load()
        prec = a.precision or w.precision
    if a.bits:
        bits = (bits * (a.bits // len(bits) + 1))[:a.bits]
    lib = _native.lib()
    lib.qx_sim_launches.restype = C.c_int64
    batch = a.batch or min(256, len(bits))
    ex = Executor(dsl, data, precision=prec, batch=batch, fuse_slices=a.fuse_slices, autotune=a.autotune,
                  graph=a.graph, kernel=a.kernel)
    print(f"finalized {a.workload}: {ex.num_slices} slices, precision {prec}, batch {batch}", flush=True)
    for i in range(a.calls):
        t0 = time.time()
        ex.amplitudes(bits)
        print(f"call {i + 1}: {len(bits)} bitstrings in {time.time() - t0:.1f} s, "
              f"{lib.qx_sim_launches()} launches checked, last call {ex.stats()['last_call_launches']} launches",
              flush=True)
    if a.vjp:
        import numpy as np

        amps = ex.amplitudes(bits)
        ex.vjp(bits, np.ones_like(amps))
        print(f"vjp: {lib.qx_sim_launches()} launches checked", flush=True)
    if a.noalloc:
        import numpy as np

        f = lib.qx_amplitudes_noalloc
        f.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
        f.restype = C.c_int32
        for nb in (len(bits), max(1, len(bits) // 3)):
            out = np.zeros(nb * ex.plan.result_elems, dtype=np.complex128)
            st = f(ex.plan._h, ex.plan._bits(bits[:nb]), nb, 0, ex.num_slices,
                   out.ctypes.data_as(C.POINTER(C.c_double)))
            if st != 0:
                raise SystemExit(f"noalloc call failed: {lib.qx_last_error().decode()}")
        print(f"noalloc: {lib.qx_sim_launches()} launches checked", flush=True)
    ex.close()
    print("sequence done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
