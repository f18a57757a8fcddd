"""Islands and sleeping on the device (PhysicsWorld::createIslands / updateSleepingBodies,
PhysicsWorld.cpp:695-926) against the reference with sleeping enabled: piles that fall asleep, a
ball that wakes the island it lands on, a hanging chain woken through its joints, and a batch of
worlds that fall asleep at different times.  Every step compares the oriented pair sets, contacts,
body states and which bodies sleep."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _run(scenes, steps, solver, tol=0.0):
    """Steps the worlds as one batch and each world's reference fed its share of the device's contact
    and joint order.  Body states must agree to `tol` (0: bit for bit) and the sleeping flags exactly.
    Returns the batch and the number of (world, step) pairs with sleeping bodies seen."""
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(scenes, solver=solver)
    refs = [RefWorld(sc) for sc in scenes]
    order = dev.joint_order()
    for w, ref in enumerate(refs):
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        ref.set_joint_order(order[(order >= j0) & (order < j1)] - j0)
    asleep_seen, woke_seen = 0, 0
    prev = [np.zeros(sc.nb_bodies, np.int32) for sc in scenes]
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        p1, p2 = dev.pairs()
        pos, q, v, w = dev.bodies()
        sl = dev.sleeping()
        for wi, ref in enumerate(refs):
            c0, c1 = dev.collider_base[wi], dev.collider_base[wi + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            local = {key: a[sel] for key, a in ct.items()}
            local["colliders"] = local["colliders"] - c0
            ref.step_ordered(DT, local["colliders"][:, 0], local["colliders"][:, 1])
            psel = (p1 >= c0) & (p1 < c1)
            r1, r2 = ref.pairs()
            assert pair_set(p1[psel] - c0, p2[psel] - c0) == pair_set(r1, r2), f"step {k} world {wi}: pair sets differ"
            errs = compare_contacts(local, ref.contacts())
            assert not errs, f"step {k} world {wi}: " + "; ".join(errs[:5])
            b0, b1 = dev.body_base[wi], dev.body_base[wi + 1]
            ref_sl = ref.sleeping()
            assert np.array_equal(sl[b0:b1], ref_sl), f"step {k} world {wi}: sleeping {sl[b0:b1]} vs {ref_sl}"
            dev_state = np.concatenate([a[b0:b1].ravel() for a in (pos, q, v, w)])
            ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
            err = float(np.max(np.abs(dev_state - ref_state)))
            assert err <= tol, f"step {k} world {wi}: body state differs by {err}"
            asleep_seen += int(ref_sl.any())
            woke_seen += int(np.any(prev[wi] & ~ref_sl))
            prev[wi] = ref_sl
    return dev, asleep_seen, woke_seen


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_pile_falls_asleep(solver):
    dev, asleep, woke = _run([S.sleeping_pile(3, 2, seed=5)], 200, solver)
    assert asleep > 30
    assert dev.sleeping().sum() == 18


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_ball_wakes_the_sleeping_pile(solver):
    dev, asleep, woke = _run([S.sleeping_pile(3, 2, seed=5, ball_height=40.0)], 300, solver)
    assert asleep > 30 and woke >= 1


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_chain_woken_through_its_joints(solver):
    dev, asleep, woke = _run([S.sleeping_chain()], 160, solver)
    assert asleep > 10 and woke >= 1


def test_batch_of_worlds_sleeping_at_different_times():
    scenes = [S.sleeping_pile(2, 2, seed=9), S.sleeping_pile(3, 2, seed=5, ball_height=40.0), S.sleeping_chain(),
              S.sleeping_pile(3, 1, seed=12), S.sleeping_pile(2, 3, seed=13)]
    dev, asleep, woke = _run(scenes, 240, "fused")
    assert asleep > 100 and woke >= 2
