"""Builds libcompas_amd.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m compas_amd.build   (or build() from __graft_entry__)
The library lands in compas_amd/libcompas_amd.so; objects in compas_amd/_build/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcompas_amd.so")

ARCH = os.environ.get("CMP_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops keeps the compiler from forming v_pk_*_f32 (all arithmetic here is FP64 anyway)
HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    units = [
        ("cmp_kernels.hip", "k2.o", ["-DCMP_KERNEL_DIM=2"]),
        ("cmp_kernels.hip", "k3.o", ["-DCMP_KERNEL_DIM=3"]),
        ("cmp_kernels.hip", "l2.o", ["-DCMP_KERNEL_DIM=2", "-DCMP_LIST_UNIT"]),
        ("cmp_kernels.hip", "l3.o", ["-DCMP_KERNEL_DIM=3", "-DCMP_LIST_UNIT"]),
        ("cmp_pf.hip", "pf2.o", ["-DCMP_KERNEL_DIM=2"]),
        ("cmp_pf.hip", "pf3.o", ["-DCMP_KERNEL_DIM=3"]),
        ("cmp_mg.hip", "mg2.o", ["-DCMP_KERNEL_DIM=2"]),
        ("cmp_mg.hip", "mg3.o", ["-DCMP_KERNEL_DIM=3"]),
        ("cmp_amr_pf.hip", "ap2.o", ["-DCMP_KERNEL_DIM=2"]),
        ("cmp_amr_pf.hip", "ap3.o", ["-DCMP_KERNEL_DIM=3"]),
        ("cmp_capi.hip", "capi.o", []),
        ("cmp_batch.hip", "batch.o", []),
        ("cmp_amr.hip", "amr.o", []),
        ("cmp_tune.hip", "tune.o", []),
        ("cmp_cluster.hip", "cluster.o", []),
    ]
    procs = []
    for src, obj, extra in units:
        cmd = [hipcc] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out.decode(errors="replace"))
            sys.stderr.write("\nFAILED: " + " ".join(cmd) + "\n")
        elif verbose and out:
            sys.stdout.write(out.decode(errors="replace"))
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    link = [hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [
        os.path.join(BUILD, o) for _, o, _ in units
    ]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.check_call(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
