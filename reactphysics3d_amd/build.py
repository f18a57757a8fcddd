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

SOURCES = ["world.cpp", "quickhull.cpp", "concave.cpp", "broadphase.hip", "narrowphase.hip", "solver.hip", "joints.hip", "fused.hip", "mlp.hip",
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


# AddressSanitizer build of the host code (C-ABI, world description, QuickHull, batch set-up); the
# device code is built as usual.  The library needs the ASan runtime preloaded; see asan_run().
ASAN_DIR = os.path.join(HERE, "_lib_asan")
ASAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def asan_runtime():
    """Path of the compiler's shared ASan runtime (clang's, as hipcc compiles the host code)."""
    out = subprocess.run(["hipcc", "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        raise RuntimeError("clang ASan runtime not found")
    return path


def asan_run(args):
    """Runs a command (e.g. the CPU test suite) against the ASan build of the library."""
    env = dict(os.environ, LD_PRELOAD=asan_runtime(), RP3DG_LIBRARY=os.path.join(ASAN_DIR, "librp3d_mi355x.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
    return subprocess.run(args, env=env).returncode


def build(verbose=False, jobs=None, asan=False):
    lib_dir = ASAN_DIR if asan else LIB_DIR
    lib = os.path.join(lib_dir, "librp3d_mi355x.so")
    extra = ASAN_FLAGS if asan else []
    os.makedirs(lib_dir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(lib_dir, src + ".o")
        cmd = ["hipcc", *HIP_FLAGS, *extra, "-c", path, "-o", obj]
        if src.endswith(".cpp"):  # host code, same FP flags
            cmd = ["hipcc", "-x", "hip", *HIP_FLAGS, *extra, "-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        list(ex.map(run, cmds))
    tmp = lib + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    if asan:
        return lib
    # the C++ API examples (host code only: g++ against the C++ mirror header and the library)
    for name in ("hello_world", "joints", "contacts", "driven", "meshes", "queries", "concave"):
        ex = os.path.join(LIB_DIR, name)
        cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(os.path.dirname(HERE), "include"),
               os.path.join(HERE, "examples", name + ".cpp"), "-L", LIB_DIR, "-lrp3d_mi355x",
               "-Wl,-rpath,$ORIGIN", "-o", ex]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    # python -m reactphysics3d_amd.build [-v] [--asan [-- command...]]: with --asan, builds the ASan
    # library and runs the command (default: the CPU test suite) against it
    if "--asan" in sys.argv:
        build(verbose="-v" in sys.argv, asan=True)
        rest = sys.argv[sys.argv.index("--") + 1:] if "--" in sys.argv else \
            [sys.executable, "-m", "pytest", "tests", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
        sys.exit(asan_run(rest))
    build(verbose="-v" in sys.argv)
    print(LIB)
