"""Device step vs the reference on a box pile (BoxShape SAT, manifold reduction, warm start,
coloured contact solver, integration, broadphase refit)."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu


def _run(sc, steps, ref_available, solver="auto"):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(sc, solver=solver)
    ref = RefWorld(sc)
    dt = 1.0 / 60.0
    worst = dict(pos=0.0, vel=0.0)
    for k in range(steps):
        dev.step(dt)
        ct = dev.contacts()
        ref.step_ordered(dt, ct["colliders"][:, 0], ct["colliders"][:, 1])
        d1, d2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert pair_set(d1, d2) == pair_set(r1, r2), f"step {k}: oriented pair sets differ"
        errs = compare_contacts(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        pd, qd, vd, wd = dev.bodies()
        pr, qr, vr, wr = ref.bodies()
        worst["pos"] = max(worst["pos"], float(np.max(np.abs(pd - pr))), float(np.max(np.abs(qd - qr))))
        worst["vel"] = max(worst["vel"], float(np.max(np.abs(vd - vr))), float(np.max(np.abs(wd - wr))))
        assert worst["pos"] <= 1e-4 and worst["vel"] <= 1e-4, f"step {k}: body state differs {worst}"
    return dev, ref, worst


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_box_pile_small(ref_available, solver):
    sc = S.box_pile(4, 3, seed=11)
    dev, ref, worst = _run(sc, 120, ref_available, solver=solver)
    print("worst", worst, dev.counts())
