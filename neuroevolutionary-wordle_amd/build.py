"""Builds the HIP runtime library libne_wordle.so for gfx950, in-tree, with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libne_wordle.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed fp32 instructions (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep fp32 denormals (the fitness floor is the smallest denormal) and single-rounding ops
    "-fno-gpu-flush-denormals-to-zero",
    "-ffp-contract=off",
    "-Wno-unused-result",
    "-Wno-cuda-compat",  # device lambdas carry __device__ after the parameter list
]

SOURCES = ["ne_kernels.hip", "ne_surrogate.hip", "ne_telemetry.hip", "ne_runtime.hip", "ne_model.hip",
           "ne_checkpoint.cpp", "ne_artifact.cpp"]


DRIVER = os.path.join(HERE, "run_ga")


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libne_wordle.so")
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


def build_driver(verbose=False):
    """The C++ GA driver over the host API (host/slab_runtime.hpp), linked to libne_wordle.so."""
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(HERE, "host", "run_ga.cpp"), "-o", DRIVER + ".tmp",
           "-L" + HERE, "-l:libne_wordle.so", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building run_ga")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
