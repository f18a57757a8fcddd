"""The C++ mirror of the reference API (include/reactphysics3d_mi355x.hpp): the helloworld example
builds against it, constructs its world through the C-ABI, and (GPU) steps it like the reference."""
import os
import subprocess

import numpy as np
import pytest

from reactphysics3d_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "reactphysics3d_amd", "_lib", "hello_world")


def _run(steps):
    if not os.path.exists(EXE):
        pytest.skip("example not built (python -m reactphysics3d_amd.build)")
    out = subprocess.run([EXE, str(steps)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [list(map(float, line.split()[2:])) for line in out.stdout.splitlines() if line.startswith("body")]
    return np.array(rows, np.float32)


def _scene():
    sc = S.Scene("hello_world", S.make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (10.0, 10.0, 0.5), friction=0.5, bounciness=0.0)
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 2.0), (0.0998334, 0.0, 0.0, 0.9950042))
    sc.add_collider(b, S.BOX, (0.5, 0.5, 0.5), friction=0.5, bounciness=0.0, density=1.0)
    s = sc.add_body(S.DYNAMIC, (3.5, 0.0, 3.0), lin=(-1.0, 0.0, 0.0))
    sc.add_collider(s, S.SPHERE, (0.4,), friction=0.5, bounciness=0.0, density=1.0)
    return sc


def test_example_builds_and_creates_world():
    st = _run(0)
    assert st.shape == (3, 13)
    assert np.allclose(st[1, :7], [0, 0, 2, 0.0998334, 0, 0, 0.9950042], atol=1e-7)
    assert np.allclose(st[2, 7:10], [-1, 0, 0])


@pytest.mark.gpu
def test_example_steps_like_the_reference(ref_available):
    from oracle.ref import RefWorld
    st = _run(120)
    ref = RefWorld(_scene())
    for _ in range(120):
        ref.step(1.0 / 60.0)   # box and ball never touch: their manifolds share only the static ground
    p, q, v, w = ref.bodies()
    want = np.concatenate([p, q, v, w], 1)
    err = np.max(np.abs(st - want))
    print("max |device - reference|", err)
    assert err <= 1e-4
