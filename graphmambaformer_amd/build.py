"""Build libgmfa.so (HIP, gfx950) in-tree with hipcc.

    python -m graphmambaformer_amd.build [--jobs N]

Each ``csrc/*.hip`` is compiled to an object in ``csrc/_build`` (in parallel),
then linked into ``graphmambaformer_amd/libgmfa.so`` against the HIP runtime.
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libgmfa.so")
# the forward-kernel variants first: the largest translation units start first
SOURCES = ["sw_tiles_24.hip", "sw_tiles_wide.hip", "sw_tiles_mid.hip", "sw_tiles_small.hip",
           "sw_split.hip", "sw.hip", "api.hip", "chain.hip", "chain_post.hip", "serve.hip",
           "seed.hip", "wfa.hip", "noalloc.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libgmfa)")


def compile_flags():
    return [
        "-O3",
        "-std=c++17",
        "-fPIC",
        f"--offload-arch={ARCH}",
        "-mcode-object-version=5",
        # no packed-FP32 instruction selection (the v_pk_*_f32 forms are kept out)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fno-gpu-rdc",
        "-I", os.path.join(ROOT, "include"),
        "-I", CSRC,
        "-Wall",
        "-Wno-unused-lambda-capture",
    ]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + compile_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 0, verbose: bool = False) -> str:
    """jobs: parallel compiles (0: one per host core, at most 16)."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(os.cpu_count() or 4, 16)
    with ThreadPoolExecutor(max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    args = ap.parse_args()
    build(args.jobs, verbose=True)
