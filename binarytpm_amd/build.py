"""Build libbtpm.so in-tree with hipcc for gfx950 (no JIT, no torch).

Usage: python -m binarytpm_amd.build   (or binarytpm_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["conduct.hip", "aux.hip", "stepwise.hip", "geometry.hip", "viewfactor.hip", "observe.hip", "io.hip", "probe.hip", "host.hip", "compat.hip"]
LIB = os.path.join(HERE, "libbtpm.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 (v_pk_*_f32) code generation: some of its forms have given
    # sporadic wrong results on this target, so every kernel is built without it
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I", CSRC, "-I", os.path.join(HERE, "..", "include"),
           *srcs, "-o", tmp, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
