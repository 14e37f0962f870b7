"""Builds the CPU oracle (liboracle.so) and, when the reference sources are present, the
reference-header probe library under oracle/_ref/.  Test infrastructure only."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_SRC = "/root/reference/src"
ORACLE_FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-pthread"]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_oracle():
    out = os.path.join(HERE, "liboracle.so")
    _run(["g++"] + ORACLE_FLAGS + [os.path.join(HERE, "ne_oracle.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_reference_probe():
    """Compiles oracle/ref_probe.cpp against the reference's integer-only headers (wordle rules,
    hint grids, toroidal grid, shard schedule), which build with g++ as they are."""
    if not os.path.isdir(REFERENCE_SRC):
        return None
    ref_dir = os.path.join(HERE, "_ref")
    os.makedirs(ref_dir, exist_ok=True)
    out = os.path.join(ref_dir, "libneref.so")
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-I", REFERENCE_SRC,
          os.path.join(HERE, "ref_probe.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_oracle())
    print(build_reference_probe())
