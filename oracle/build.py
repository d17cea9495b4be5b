"""Builds the oracle (CPU restatement, test infrastructure) as oracle/_build/libcpu_ref.so.

gcc, C99, -O3, no FMA contraction (-ffp-contract=off), no fast-math, OpenMP for the timed CPU
baseline. build_variant() compiles the same source with other flags elsewhere (the bench's
-march=native baseline, built on the host that runs it). The reference itself (COMPAS) cannot be built here: it needs AMReX 26.01, which is not
vendored and not installed, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libcpu_ref.so")

CFLAGS = ["-O3", "-std=c99", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + CFLAGS + ["-shared", "-o", tmp, os.path.join(HERE, "cpu_ref.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_variant(out_path, extra_flags, verbose=False):
    """The oracle built with CFLAGS + extra_flags into out_path (a temporary location)."""
    cmd = ["gcc"] + CFLAGS + list(extra_flags) + ["-shared", "-o", out_path, os.path.join(HERE, "cpu_ref.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out_path


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
