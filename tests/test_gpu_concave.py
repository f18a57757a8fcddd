"""Convex shapes on concave ones on the device against the reference (CollisionDetectionSystem::
computeConvexVsConcaveMiddlePhase, CD.cpp:522-615; the triangle tests with TriangleShape and smooth mesh
contacts, TriangleShape.cpp:85-114; processPotentialContacts / reducePotentialContactManifolds for pairs with
several manifolds, CD.cpp:1292-1458; warm starting over a pair's manifolds, :1069-1160): spheres, capsules,
boxes and convex meshes dropped on a triangle mesh (computed or given vertex normals, scaled) and on a
height field, stepped with the device's contact order fed to the oracle.  Pair sets and per-manifold point
counts are equal, contacts within 1e-5 and bodies bit for bit, every step."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _run(sc, steps, solver="auto"):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(sc, solver=solver)
    ref = RefWorld(sc)
    extra = 0
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        ref.step_ordered(DT, ct["colliders"][:, 0], ct["colliders"][:, 1], ct["manifold"])
        d1, d2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert pair_set(d1, d2) == pair_set(r1, r2), f"step {k}: pair sets differ"
        errs = compare_contacts(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        dev_state = np.concatenate([a.ravel() for a in dev.bodies()])
        ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
        err = float(np.max(np.abs(dev_state - ref_state)))
        assert err == 0.0, f"step {k}: body state differs by {err}"
        extra += int(np.sum(ct["manifold"] > 0))
    return dev, extra


SHAPES = {"spheres": (S.SPHERE,), "capsules": (S.CAPSULE,), "boxes": (S.BOX,), "hulls": (S.HULL,),
          "mixed": (S.SPHERE, S.CAPSULE, S.BOX, S.HULL)}


@pytest.mark.parametrize("shapes", list(SHAPES))
@pytest.mark.parametrize("kind", ["mesh", "heightfield"])
def test_convex_shapes_on_concave_ground(kind, shapes):
    sc = S.concave_scene(kind, n_bodies=24, shapes=SHAPES[shapes], seed=9)
    dev, extra = _run(sc, 120)
    print(kind, shapes, "extra manifolds", extra, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_scaled_mesh_with_given_normals_both_layouts(solver):
    sc = S.concave_scene("mesh", n_bodies=30, scale=(1.3, 0.8, 1.1), normals=True, seed=11)
    dev, extra = _run(sc, 120, solver=solver)
    print("extra manifolds", extra)


def test_scaled_heightfield():
    sc = S.concave_scene("heightfield", n_bodies=30, scale=(0.9, 1.4, 1.2), seed=12)
    dev, extra = _run(sc, 120)
    print("extra manifolds", extra)
