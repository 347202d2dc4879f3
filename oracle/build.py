"""Build the CPU oracle (test infrastructure only) into oracle/liboracle.so.

The oracle is a plain C++ restatement of the reference; it is compiled with
g++ and -ffp-contract=off so that its floating-point operation sequence is
exactly what the source says. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")
SOURCES = ["oracle_hotpath.cpp", "oracle_math.cpp", "oracle_bins.cpp", "oracle_hull_fbp.cpp", "oracle_net.cpp"]
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp"]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = ["g++", *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the oracle")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
