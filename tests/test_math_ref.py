"""The device's acos / atan2 restatement (rp3d_math.h ref_acosf / ref_atan2f, fdlibm's float method)
returns the same bits as the C library functions the reference build calls.  Host build of the same
header with hipcc; the device build of the same code is covered by the bit-exact hinge-limit and
cone-limit parity scenes."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not available")
def test_ref_acos_atan2_match_the_c_library(tmp_path):
    exe = str(tmp_path / "math_ref_check")
    subprocess.run(["hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-Wno-unused-result", "-I", os.path.join(ROOT, "reactphysics3d_amd", "csrc"),
                    os.path.join(HERE, "native", "math_ref_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    # stride 61 keeps the run to a few seconds: ~35M acos, ~5M atan and ~0.7M atan2 arguments
    r = subprocess.run([exe, "61"], capture_output=True, text=True)
    print(r.stdout)
    counts = {line.split()[0]: (int(line.split()[1]), int(line.split()[2])) for line in r.stdout.split("\n") if line}
    for name in ("acosf", "atanf", "atan2f"):
        bad, n = counts[name]
        assert n > 100000 and bad == 0, f"{name}: {bad} of {n} differ from the C library"
    assert r.returncode == 0
