"""Build libqxhip.so for gfx950: host C++ with g++, kernels with hipcc, linked by hipcc.

Usage: python qxtools.jl_amd/build.py [--debug]
The library is written to qxtools.jl_amd/qxhip/libqxhip.so (in-tree)."""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "qxhip", "libqxhip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_SRC = ["dsl.cpp", "lower.cpp", "runtime.cpp"]
HIP_SRC = ["kernels.hip"]
# -amdgpu-mfma-vgpr-form: matrix-core products accumulate in ordinary VGPRs; with
# accumulation registers, the K loops of the matrix-core GEMM copied every loop-carried
# accumulator between the two register files on each trip
HIPFLAGS = ["--offload-arch=gfx950", "-std=c++17", "-fPIC", "-Xclang", "-target-feature",
            "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(debug: bool = False, verbose: bool = False) -> str:
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    opt = ["-O0", "-g"] if debug else ["-O3"]
    jobs = []
    for f in HOST_SRC:
        o = os.path.join(bdir, f + ".o")
        jobs.append(["g++", "-std=c++17", "-fPIC", "-Wall", "-Wextra", *opt, "-D__HIP_PLATFORM_AMD__",
                     f"-I{ROCM}/include", "-c", os.path.join(CSRC, f), "-o", o])
    for f in HIP_SRC:
        o = os.path.join(bdir, f + ".o")
        jobs.append([f"{ROCM}/bin/hipcc", *HIPFLAGS, *opt, "-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    objs = [os.path.join(bdir, f + ".o") for f in HOST_SRC + HIP_SRC]
    tmp = OUT + ".tmp"
    _run([f"{ROCM}/bin/hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
          f"-L{ROCM}/lib", "-lrocblas", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(debug="--debug" in sys.argv, verbose=True))
