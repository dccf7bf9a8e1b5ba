"""The host C++ (DSL parser, lowering, runtime) built with AddressSanitizer and
UndefinedBehaviorSanitizer against a simulated device (qxtools.jl_amd/sanitize/):
device memory is host memory, and every kernel launch and library GEMM checks, from
its arguments alone, that all elements it may touch lie inside one live device
allocation -- eagerly and on every replay of a captured graph.  Call sequences of the
executor's paths run under it: autotuning, graph capture and replay, looped slices,
the pullback, the allocation-free entry point, and the 7x7 workload with all 1024
slices looped (about 150K launches per call), with the tuning call first."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "qxtools.jl_amd", "sanitize")


@pytest.fixture(scope="module")
def san_env():
    sys.path.insert(0, SAN)
    import build_sanitized

    lib = build_sanitized.build()
    rt = " ".join(subprocess.run(["g++", f"-print-file-name={n}"], capture_output=True, text=True,
                                 check=True).stdout.strip() for n in ("libasan.so", "libstdc++.so"))
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1", QXHIP_LIB=lib)
    return env


def _run(env, *args):
    r = subprocess.run([sys.executable, os.path.join(SAN, "run_sequence.py"), *args], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "sequence done" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "VIOLATION" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    return r.stdout


@pytest.mark.parametrize("args", [
    ("rqc_3x3_d16", "--autotune", "1", "--graph", "1", "--calls", "3", "--vjp", "--noalloc"),
    ("rqc_5x5_d20", "--autotune", "1", "--graph", "1", "--calls", "2", "--vjp"),
    ("sliced", "--fuse-slices", "0", "--graph", "1", "--calls", "3", "--vjp", "--noalloc", "--batch", "4"),
    ("sliced", "--precision", "c64_bf16x3", "--autotune", "1", "--graph", "1"),
    ("rqc_6x6_d12_s6", "--autotune", "1", "--graph", "1", "--bits", "64"),
    ("rqc_7x7_d12_s10_open5", "--fuse-slices", "0", "--autotune", "1", "--graph", "1", "--calls", "3"),
    ("plain:0,4,4,12,3", "--kernel", "11", "--graph", "1"),
    ("plain:2,6,5,7,5", "--kernel", "11", "--precision", "c64_3m"),
    ("plain:1,7,6,8,6", "--kernel", "11", "--precision", "c64_4m", "--autotune", "1"),
], ids=["config1", "config2", "sliced-looped", "sliced-bf16", "6x6", "7x7-looped", "plain-splitk", "plain-3m",
        "plain-4m-tuned"])
def test_call_sequences_under_sanitizers(san_env, args):
    out = _run(san_env, *args)
    assert "launches checked" in out
