"""Builds the HIP step library (librp3d_mi355x.so) in-tree for gfx950.

Run as ``python -m reactphysics3d_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "librp3d_mi355x.so")

SOURCES = ["world.cpp", "quickhull.cpp", "broadphase.hip", "narrowphase.hip", "solver.hip", "joints.hip", "fused.hip", "mlp.hip",
           "islands.hip", "events.hip", "raycast.hip", "api.hip",
           "batch.hip"]

# -ffp-contract=off + IEEE div/sqrt keep the FP32 operation sequence of the x86-64 reference build;
# packed FP32 instructions are disabled (wrong results have been seen from some of their forms).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False, jobs=None):
    os.makedirs(LIB_DIR, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(LIB_DIR, src + ".o")
        cmd = ["hipcc", *HIP_FLAGS, "-c", path, "-o", obj]
        if src.endswith(".cpp"):  # host code, same FP flags
            cmd = ["hipcc", "-x", "hip", *HIP_FLAGS, "-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        list(ex.map(run, cmds))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the C++ API examples (host code only: g++ against the C++ mirror header and the library)
    for name in ("hello_world", "joints", "contacts", "driven", "meshes", "queries"):
        ex = os.path.join(LIB_DIR, name)
        cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(os.path.dirname(HERE), "include"),
               os.path.join(HERE, "examples", name + ".cpp"), "-L", LIB_DIR, "-lrp3d_mi355x",
               "-Wl,-rpath,$ORIGIN", "-o", ex]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
