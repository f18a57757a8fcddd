"""Device step vs the reference on mixed spheres / boxes / capsules in a walled container:
every narrowphase pair type (closed forms, GJK, SAT fallbacks, box SAT)."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu


def _run(sc, steps, tol_state=1e-4, solver="auto"):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(sc, solver=solver)
    ref = RefWorld(sc)
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
        assert worst <= tol_state, f"step {k}: body state differs by {worst}"
    return dev, worst


@pytest.mark.parametrize("seed,solver", [(2, "fused"), (5, "per_color")])
def test_mixed_container_small(seed, solver):
    sc = S.mixed_container(600, extent=(24.0, 24.0, 12.0), seed=seed)
    dev, worst = _run(sc, 90, solver=solver)
    print("worst", worst, dev.counts())
