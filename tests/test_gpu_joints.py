"""Device joint solvers (ball-and-socket, fixed, hinge; velocity + non-linear Gauss-Seidel position
correction) vs the reference fed the device's joint and contact order."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu


def _run(sc, steps, tol=1e-4, solver="auto"):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(sc, solver=solver)
    ref = RefWorld(sc)
    ref.set_joint_order(dev.joint_order())
    dt = 1.0 / 60.0
    worst = 0.0
    for k in range(steps):
        dev.step(dt)
        ct = dev.contacts()
        ref.step_ordered(dt, ct["colliders"][:, 0], ct["colliders"][:, 1])
        d1, d2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert pair_set(d1, d2) == pair_set(r1, r2), f"step {k}: oriented pair sets differ"
        errs = compare_contacts(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        dev_state = np.concatenate([a.ravel() for a in dev.bodies()])
        ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
        worst = max(worst, float(np.max(np.abs(dev_state - ref_state))))
        assert worst <= tol, f"step {k}: body state differs by {worst}"
    return dev, worst


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_ball_socket_chains(solver):
    sc = S.sphere_chains(n_chains=9, length=8, seed=31)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_hinge_fixed_tower(solver):
    sc = S.jointed_tower(height=8, seed=41)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())
