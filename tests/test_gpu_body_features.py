"""Kinematic bodies, damping, lock factors, disabled gravity and external forces / torques applied
between steps (RigidBody::applyWorldForceAtCenterOfMass / applyWorldTorque), on the device against
the reference fed the device's contact order: pair sets, contacts and body states bit for bit."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_body_features_with_forces(solver):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    sc = S.body_features()
    dev = Batch([sc], solver=solver)
    ref = RefWorld(sc)
    rng = np.random.default_rng(5)
    nb = sc.nb_bodies
    for k in range(150):
        if k % 3 == 0:
            f = rng.normal(0.0, 20.0, (nb, 3)).astype(np.float32)
            t = rng.normal(0.0, 2.0, (nb, 3)).astype(np.float32)
            f[rng.random(nb) < 0.5] = 0.0
            dev.apply_forces(f, t)
            ref.apply_forces(f, t)
        dev.step(DT)
        ct = dev.contacts()
        ref.step_ordered(DT, ct["colliders"][:, 0], ct["colliders"][:, 1])
        d1, d2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert pair_set(d1, d2) == pair_set(r1, r2), f"step {k}: pair sets differ"
        errs = compare_contacts(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        dev_state = np.concatenate([a.ravel() for a in dev.bodies()])
        ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
        err = float(np.max(np.abs(dev_state - ref_state)))
        assert err == 0.0, f"step {k}: body state differs by {err}"
    p, q, v, w = dev.bodies()
    # the kinematic platform kept its velocities; lock factors only scale changes, so the locked box kept
    # its initial y velocity and its initial x and z spin through contacts and random forces
    assert np.array_equal(v[1], np.float32([0.5, 0.0, 0.0])) and np.array_equal(w[1], np.float32([0.0, 0.0, 0.3]))
    assert v[7, 1] == 1.0 and w[7, 0] == 2.0 and w[7, 2] == 3.0
