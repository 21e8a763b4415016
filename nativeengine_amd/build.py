"""Build step for the HIP library (gfx950) -- in-tree, no JIT cache.

    python -m nativeengine_amd.build

compiles nativeengine_amd/csrc/ne_gpu.hip into nativeengine_amd/libnegpu.so
with hipcc for gfx950 only.  Floating point is IEEE double with no fused
multiply-add contraction (-ffp-contract=off): the kernels are meant to give
the reference's bit patterns.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnegpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "ne_gpu.hip")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for ne_gpu.hip")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
