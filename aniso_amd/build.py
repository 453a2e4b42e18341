"""Builds libaniso_amd.so in-tree with hipcc for gfx950.

    python -m aniso_amd.build

Flags:
* ``-ffp-contract=off``: the preprocess must round every multiply and add separately so its
  splat records, keys and tile rects match the CPU oracle bit for bit; kernels that want fused
  multiply-adds write them explicitly.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: keep the compiler from emitting packed
  FP32 instructions.
* HIP's default correctly rounded FP32 division and square root are kept (no fast-math).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaniso_amd.so")
SOURCES = ["renderer.hip", "diff.hip", "batch.hip", "diffviews.hip", "noalloc.hip", "ingest.hip", "io.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


def command(out: str = LIB):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [HIPCC, "-O3", "-std=c++20", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, *srcs, "-o", out]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libaniso_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
