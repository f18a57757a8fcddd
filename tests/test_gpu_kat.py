"""The device narrowphase against the reference's own known-answer contact points, and a closed-form
check of setting body state between steps."""
import pytest

from tests import kat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", kat.load()["cases"], ids=lambda c: c["name"])
def test_device_matches_reference_kat(case):
    from reactphysics3d_amd.lib import Batch
    sc, _ = kat.scene(case)
    dev = Batch(sc)
    dev.step(1.0 / 60.0)
    problems = kat.check(case, dev.contacts(), kat.load()["epsilon"])
    assert not problems, problems


def test_set_bodies_teleport_then_free_fall():
    """rp3dg_batch_set_bodies: a body moved and stopped mid-run then falls by the symplectic-Euler
    closed form from its new position."""
    import numpy as np

    from reactphysics3d_amd import scene as S
    from reactphysics3d_amd.lib import Batch
    sc = S.Scene("teleport", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 10.0))
    sc.add_collider(b, S.SPHERE, (0.5,))
    dev = Batch(sc)
    dt = 1.0 / 60.0
    dev.step(dt, 10)
    p, q, v, w = dev.bodies()
    p[0] = (3.0, -2.0, 50.0)
    v[0] = 0.0
    dev.set_bodies(pos=p, lin=v)
    for k in range(1, 31):
        dev.step(dt)
        p, _, v, _ = dev.bodies()
        assert abs(v[0, 2] + 9.81 * k * dt) <= 1e-4 * k
        assert abs(p[0, 2] - (50.0 - 9.81 * dt * dt * k * (k + 1) / 2)) <= 2e-3
        assert abs(p[0, 0] - 3.0) <= 1e-5 and abs(p[0, 1] + 2.0) <= 1e-5
