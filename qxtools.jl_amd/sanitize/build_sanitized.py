"""Build libqxhip's host code with AddressSanitizer and UndefinedBehaviorSanitizer
against the simulated device of this folder (no GPU, no HIP runtime, no rocBLAS):
every kernel launch and library GEMM is checked for the reach of its operands instead
of being run (sim_device.cpp).

    python qxtools.jl_amd/sanitize/build_sanitized.py     # -> qxtools.jl_amd/_build/libqxhip_san.so

Load it with the ASan runtime preloaded, e.g.
    LD_PRELOAD=$(g++ -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \\
    QXHIP_LIB=qxtools.jl_amd/_build/libqxhip_san.so python ...
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "_build", "libqxhip_san.so")
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("dsl.cpp", "lower.cpp", "runtime.cpp")] + \
          [os.path.join(HERE, "sim_device.cpp")]
FLAGS = ["-std=c++17", "-fPIC", "-shared", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-Wall", "-Wextra"]


def asan_runtime() -> str:
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", *FLAGS, f"-I{os.path.join(HERE, 'include')}", *SOURCES, "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitizer build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
