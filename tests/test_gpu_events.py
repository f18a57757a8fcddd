"""Contact events (EventListener::onContact, CollisionDetectionSystem::reportContacts,
CollisionDetectionSystem.cpp:1708-1741) on the device against the reference's listener: every step
the sets of (collider1, collider2, ContactStart / ContactStay / ContactExit, contact points) agree,
for a pile, a bouncing mixed container (pairs removed while colliding), a pile woken from sleep (pairs
re-enabled with their colliding flags) and a batch of worlds."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _run(scenes, steps, solver="auto", graph_every=0):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(scenes, solver=solver)
    dev.set_contact_events(True)
    refs = [RefWorld(sc) for sc in scenes]
    for r in refs:
        r.enable_events()
    order = dev.joint_order()
    for w, ref in enumerate(refs):
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        ref.set_joint_order(order[(order >= j0) & (order < j1)] - j0)
    seen = np.zeros(3, np.int64)
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        ev = dev.contact_events()
        for wi, ref in enumerate(refs):
            c0, c1 = dev.collider_base[wi], dev.collider_base[wi + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            ref.step_ordered(DT, ct["colliders"][sel, 0] - c0, ct["colliders"][sel, 1] - c0)
            mine = ev[(ev[:, 0] >= c0) & (ev[:, 0] < c1)].copy()
            mine[:, :2] -= c0
            want = ref.events()
            assert np.array_equal(mine, want), f"step {k} world {wi}:\n{mine}\nvs reference\n{want}"
            seen += np.bincount(want[:, 2], minlength=3)
    return dev, seen


def test_pile_events():
    dev, seen = _run([S.box_pile(3, 2, seed=21)], 120)
    print("start / stay / exit", seen)
    assert seen[0] > 0 and seen[1] > 0


def test_bouncing_container_events_with_exits():
    dev, seen = _run([S.mixed_container(300, extent=(12.0, 12.0, 8.0), seed=3)], 90)
    print("start / stay / exit", seen)
    assert seen[2] > 0


def test_events_across_sleep_and_wake():
    dev, seen = _run([S.sleeping_pile(3, 2, seed=5, ball_height=40.0)], 260)
    print("start / stay / exit", seen)
    assert seen.min() > 0


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_batched_world_events(solver):
    scenes = [S.batched_world(w, seed=100, min_bodies=8, max_bodies=48, max_joints=8) for w in range(8)]
    dev, seen = _run(scenes, 60, solver)
    print("start / stay / exit", seen)
    assert seen[0] > 0 and seen[1] > 0
