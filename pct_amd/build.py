"""Build libpct.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Every source is compiled to an object in parallel (into pct_amd/_build/), then
linked into pct_amd/libpct.so.

Usage: python -m pct_amd.build   (or pct_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpct.so")
SOURCES = ["pct_ctx.hip", "pct_intersect.hip", "pct_mlp.hip", "pct_drop.hip", "pct_plan.hip", "pct_batch.hip",
           "pct_bins.hip", "pct_hull.hip", "pct_fbp.hip", "pct_net.hip", "pct_research.hip", "pct_noalloc.hip",
           "pct_io.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # identical FP results on host and device: no FMA contraction anywhere
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the dispatch preloads leading scalar kernel arguments into SGPRs: a wave
    # starts without a scalar load of its arguments (latency-bound kernels)
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-fopenmp", "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc, *HIPCC_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(SOURCES, objs)]
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed compiling libpct sources")
    tmp = OUT + ".tmp"
    if _run([hipcc, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs], verbose):
        raise RuntimeError("hipcc failed linking libpct.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
