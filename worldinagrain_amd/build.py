"""Builds libwig.so (the HIP liquid step, gfx950) in-tree with hipcc.

-ffp-contract=off: no fused multiply-adds, so each f32 operation rounds as
the reference's WGSL writes it; division and square root correctly rounded;
denormals kept; packed FP32 instructions off."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwig.so")
SOURCES = ["wig_kernels.hip", "wig_api.hip", "wig_air.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


NAPI_SRC = os.path.join(HERE, "js", "wig_napi.c")
NAPI_OUT = os.path.join(HERE, "js", "wig.node")


def build_addon(verbose=False):
    """The Node-API addon (wig.node) over libwig.so, if Node's headers are here."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I", inc, "-o", NAPI_OUT + ".tmp", NAPI_SRC,
           "-L", HERE, "-lwig", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building wig.node failed")
    os.replace(NAPI_OUT + ".tmp", NAPI_OUT)
    return NAPI_OUT


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libwig.so")
    os.replace(LIB + ".tmp", LIB)
    build_addon(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
