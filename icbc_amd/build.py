"""Builds libicbc_amd.so (HIP kernels + C ABI + icbc:: API) in-tree for gfx950.

    python -m icbc_amd.build          # or: python icbc_amd/build.py

Compiler flags that the encoder's bit-exactness depends on:
  -ffp-contract=off                    no fused multiply-add contraction
  -fno-fast-math                       IEEE semantics (NaN ordering, no reassociation)
  -fno-gpu-flush-denormals-to-zero     fp32 denormals kept, as on the host
  (HIP's default correctly-rounded fp32 division is left on)
Packed-fp32 instruction selection is disabled for device code.

Device code uses the compiler's default instruction scheduling.  The other
strategies (-mllvm -amdgpu-sched-strategy=...) were measured against it on the
current kernels: max-ilp was slower on every workload (the level-8 texture
batch by 2.5 %, level 9 by 3 %), max-memory-clause by 1 %, iterative-ilp by
2 % and iterative-minreg by 3 % (level 1 three times slower).  Scheduling
reorders instructions only and changes no arithmetic.

The library exports only the public API (csrc/exports.map).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libicbc_amd.so")
SOURCES = ["bc1_encoder.hip", "texture_prepass.hip", "transfer.hip", "icbc_runtime.cpp", "icbc_server.cpp",
           "containers.cpp", "autotune.cpp", "block_ring.cpp"]
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# Benchmark programs run by bench.py: the drop-in per-block benchmark (a C++
# caller of the icbc:: API) and the serving client (a C++ caller of the
# request server).
BENCH_SRC = os.path.join(CSRC, "drop_in_bench.cpp")
BENCH_BIN = os.path.join(HERE, "drop_in_bench")
SERVE_SRC = os.path.join(CSRC, "serve_bench.cpp")
SERVE_BIN = os.path.join(HERE, "serve_bench")


def build_drop_in_bench(verbose=False):
    for src, out in ((BENCH_SRC, BENCH_BIN), (SERVE_SRC, SERVE_BIN)):
        cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(ROOT, "include"), src,
               "-L", HERE, "-licbc_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    return BENCH_BIN


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"),
           f"-Wl,--version-script={os.path.join(CSRC, 'exports.map')}", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_drop_in_bench(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
