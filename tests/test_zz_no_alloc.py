"""qx_amplitudes_noalloc through the C ABI: the allocation-free variant of qx_amplitudes
must give that entry point's results (complex128: 1e-12 relative per amplitude, and
bit for bit where both contract with the same batch size), on two calls in a row with
the second one smaller, on a plan whose slices are looped and summed on the device."""

import ctypes as C

import numpy as np
import pytest

import qxoracle as O
import qxtest
from qxhip import Executor, _native, circuits, simulation, workloads

pytestmark = pytest.mark.gpu


def _noalloc(plan, bits, s0=0, s1=None):
    L = _native.lib()
    f = L.qx_amplitudes_noalloc
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
    f.restype = C.c_int32
    nb = len(bits)
    out = np.zeros(nb * plan.result_elems, dtype=np.complex128)
    st = f(plan._h, plan._bits(bits), nb, s0, plan.num_slices if s1 is None else s1,
           out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out


def _rel(a, r):
    return float(np.max(np.abs(a - r) / np.abs(r)))


def test_noalloc_matches_amplitudes_config1():
    dsl, data, bits = workloads.WORKLOADS["rqc_3x3_d16"].load()
    ex = Executor(dsl, data, batch=64)
    try:
        ref = ex.amplitudes(bits)                 # sizes the plan's memory for batch 64
        st, out1 = _noalloc(ex.plan, bits)
        assert st == 0, _native.lib().qx_last_error()
        assert np.array_equal(out1, ref)
        small = bits[:5]
        ref5 = ex.amplitudes(small)               # qx_amplitudes contracts these in a batch of 8
        st, out2 = _noalloc(ex.plan, small)       # batch 64, 5 valid rows
        assert st == 0, _native.lib().qx_last_error()
        assert _rel(out2, ref5) <= 1e-12
        assert np.array_equal(out2, out1[:5])
    finally:
        ex.close()


def test_noalloc_looped_slices_and_ranges():
    circ = circuits.create_rqc_circuit(2, 3, 6, seed=21)
    t, plan, g = qxtest.graph_for(circ, planner="greedy", nslice=2)
    dsl = g.to_dsl()
    bits = simulation.amplitudes_uniform(6, 99, 8)
    ex = Executor(dsl, g.data, batch=8, fuse_slices=0)
    try:
        assert ex.num_slices == 4
        ref = ex.amplitudes(bits)
        st, out1 = _noalloc(ex.plan, bits)
        assert st == 0, _native.lib().qx_last_error()
        assert np.array_equal(out1, ref)
        assert _rel(out1, O.amplitudes(dsl, g.data, bits)) <= 1e-12
        ref3 = ex.amplitudes(bits[:3], 1, 3)
        st, out2 = _noalloc(ex.plan, bits[:3], 1, 3)
        assert st == 0, _native.lib().qx_last_error()
        assert _rel(out2, ref3) <= 1e-12
        part = O.amplitudes(dsl, g.data, bits[:3], slices=list(O.slice_assignments(O.parse(dsl)))[1:3])
        assert _rel(out2, part) <= 1e-12
    finally:
        ex.close()


def test_noalloc_rejects_unsized_plan_and_oversized_call():
    dsl, data, bits = workloads.WORKLOADS["rqc_3x3_d16"].load()
    ex = Executor(dsl, data, batch=64)
    try:
        st, _ = _noalloc(ex.plan, bits[:4])       # no qx_amplitudes call yet: memory unsized
        assert st == -7
        ex.amplitudes(bits)
        many = bits * 4096                        # 262144 amplitudes: over the 8 MiB scratch
        st, _ = _noalloc(ex.plan, many)
        assert st == -1
        assert b"scratch" in _native.lib().qx_last_error()
        st, out = _noalloc(ex.plan, bits)          # still usable afterwards
        assert st == 0 and np.array_equal(out, ex.amplitudes(bits))
    finally:
        ex.close()
