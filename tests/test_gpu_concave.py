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


def _run_worlds(scenes, steps, solver="auto"):
    """Several worlds in one batch with contact events: per world, pair sets, contacts, contact and
    trigger events, bodies and sleeping states equal the reference's every step."""
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(scenes, solver=solver)
    dev.set_contact_events(True)
    refs = [RefWorld(sc) for sc in scenes]
    for r in refs:
        r.enable_events()
    seen = {"events": 0, "extra": 0, "asleep": 0}
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        p1, p2 = dev.pairs()
        ev = dev.contact_events()
        pos, q, v, w = dev.bodies()
        asleep = dev.sleeping()
        for wi, ref in enumerate(refs):
            c0, c1 = dev.collider_base[wi], dev.collider_base[wi + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            local = {key: a[sel] for key, a in ct.items()}
            local["colliders"] = local["colliders"] - c0
            ref.step_ordered(DT, local["colliders"][:, 0], local["colliders"][:, 1], local["manifold"])
            ps = (p1 >= c0) & (p1 < c1)
            r1, r2 = ref.pairs()
            assert pair_set(p1[ps] - c0, p2[ps] - c0) == pair_set(r1, r2), f"step {k} world {wi}: pair sets differ"
            errs = compare_contacts(local, ref.contacts())
            assert not errs, f"step {k} world {wi}: " + "; ".join(errs[:5])
            mine = ev[(ev[:, 0] >= c0) & (ev[:, 0] < c1)].copy()
            mine[:, :2] -= c0
            want = ref.events()
            assert np.array_equal(mine, want), f"step {k} world {wi}: events\n{mine}\nvs reference\n{want}"
            b0, b1 = dev.body_base[wi], dev.body_base[wi + 1]
            dev_state = np.concatenate([a[b0:b1].ravel() for a in (pos, q, v, w)])
            ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
            err = float(np.max(np.abs(dev_state - ref_state)))
            assert err == 0.0, f"step {k} world {wi}: body state differs by {err}"
            assert np.array_equal(asleep[b0:b1], ref.sleeping()), f"step {k} world {wi}: sleeping states differ"
            seen["events"] += len(want)
            seen["extra"] += int(np.sum(local["manifold"] > 0))
            seen["asleep"] += int(np.sum(asleep[b0:b1]))
    return dev, seen


def test_events_and_triggers_on_concave_ground():
    """ContactStart / Stay / Exit of pairs with several manifolds (points summed over them) and trigger
    overlaps with the concave ground."""
    sc = S.concave_scene("mesh", n_bodies=30, trigger_every=4, seed=21)
    dev, seen = _run_worlds([sc], 150)
    print(seen)
    assert seen["events"] > 0 and seen["extra"] > 0


def test_bodies_fall_asleep_on_concave_ground():
    """Sleeping on: bodies resting on a height field fall asleep; the pairs of disabled bodies are not
    tested and their triangles' last-frame infos start over when they wake."""
    sc = S.concave_scene("heightfield", n_bodies=20, shapes=(S.BOX, S.HULL), sleeping=True, height=(0.8, 2.0), seed=22)
    dev, seen = _run_worlds([sc], 240)
    print(seen)
    assert seen["asleep"] > 0


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_kinematic_concave_platform(solver):
    """A kinematic triangle mesh moving and turning under the falling bodies: its AABB is refit every
    step and the middle phase works in its moving frame."""
    sc = S.concave_scene("mesh", n_bodies=20, ground_type=S.KINEMATIC, ground_lin=(0.3, -0.2, 0.15),
                         ground_ang=(0.0, 0.0, 0.25), seed=23)
    dev, seen = _run_worlds([sc], 150, solver=solver)
    print(seen)


def test_batch_of_concave_worlds():
    """Uneven worlds with concave grounds in one batch (world groups of the fused layout)."""
    scenes = [S.concave_scene("mesh", n_bodies=n, nx=6 + n // 4, seed=30 + n) for n in (4, 9, 17, 30)]
    dev, seen = _run_worlds(scenes, 120)
    print(seen)
