"""Builds libvexelhip.so for gfx950 from csrc/ with hipcc, in-tree (vexhip/libvexelhip.so).

Flags that carry meaning:
  -ffp-contract=off                          no fused multiply-adds unless written as fmaf: the
                                             reference's IR never fuses, and the HLL step and the
                                             grid/G2P passes reproduce its rounding bit for bit
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
  -munsafe-fp-atomics                        hardware float atomics (global_atomic_add_f32)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                             no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "vexhip", "libvexelhip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["runtime.hip", "hll.hip", "mpm3.hip", "mpm2.hip", "sort.hip", "surrogate.hip", "fixed_scatter.hip", "diagnostics.hip", "lod.hip", "tension.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
