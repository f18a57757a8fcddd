"""Builds the HIP step library (librp3d_mi355x.so) in-tree for gfx950.

Run as ``python -m reactphysics3d_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "librp3d_mi355x.so")

SOURCES = ["world.cpp", "broadphase.hip", "narrowphase.hip", "solver.hip", "joints.hip", "batch.hip"]

# -ffp-contract=off + IEEE div/sqrt keep the FP32 operation sequence of the x86-64 reference build;
# packed FP32 instructions are disabled (wrong results have been seen from some of their forms).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(LIB_DIR, src + ".o")
        cmd = ["hipcc", *HIP_FLAGS, "-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = ["hipcc", "-x", "hip", *HIP_FLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
