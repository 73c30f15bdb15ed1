"""Build the CPU oracle ``oracle/lib/libvare_oracle.so`` and its OpenMP build
``oracle/lib/libvare_oracle_omp.so`` (test infrastructure and the CPU baseline only).

g++ -O3 with -ffp-contract=off so the FP32 instance keeps the reference's operation order.
The reference itself (CUDA 8 + thrust + OpenVDB/TBB/GL headers) cannot be compiled in this image,
so there is no ``oracle/_ref`` build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "flip_oracle.cpp")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libvare_oracle.so")
# the multi-core CPU baseline: the same source with OpenMP (parallel loops, slab-coloured P2G).
# x86-64-v3 (AVX2) rather than -march=native: the library is built on one machine and runs on another's
# host, whose CPU model may differ
OUT_OMP = os.path.join(OUT_DIR, "libvare_oracle_omp.so")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"]
FLAGS_OMP = FLAGS + ["-march=x86-64-v3", "-fopenmp"]


def _compile(flags, out, verbose):
    tmp = out + ".tmp"
    cmd = [os.environ.get("CXX", "g++")] + flags + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building %s" % out)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    _compile(FLAGS, OUT, verbose)
    _compile(FLAGS_OMP, OUT_OMP, verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
