"""Snapshot and restore of a batch (rp3dg_batch_snapshot / rp3dg_batch_restore; the reference keeps no
checkpoint, so the property checked is replay): after a restore, the steps that follow equal, bit for
bit, the steps that followed the snapshot -- bodies, sleeping states, pairs, contacts with their warm-start
impulses, contact events and the solver's per-world iteration counts -- whether they run eagerly or as
graph replays, and changes made after the snapshot (RigidBody calls, queued or applied) are undone."""
import numpy as np
import pytest

from reactphysics3d_amd import lib
from reactphysics3d_amd import scene as S

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _record(dev, events):
    pos, q, v, w = dev.bodies()
    rec = {"bodies": np.concatenate([a.ravel() for a in (pos, q, v, w)]), "sleeping": dev.sleeping()}
    p1, p2 = dev.pairs()
    rec["pairs"] = np.stack([p1, p2])
    ct = dev.contacts()
    for k, a in ct.items():
        rec["ct_" + k] = np.asarray(a)
    if events:
        rec["events"] = dev.contact_events()
    return rec


def _run(dev, steps, graph, events):
    out = []
    for k in range(steps):
        if graph and k % 2:
            dev.step_graph(DT, 1)
        else:
            dev.step(DT)
        out.append(_record(dev, events))
    return out


def _same(a, b, what):
    for k, (ra, rb) in enumerate(zip(a, b)):
        assert ra.keys() == rb.keys()
        for key in ra:
            x, y = ra[key], rb[key]
            assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), \
                f"{what}: step {k} after the snapshot, {key} differs"


def _batch(scenes, **kw):
    """One batch of the scenes; the worlds of a batch share the first scene's WorldSettings."""
    for sc in scenes[1:]:
        sc.settings = scenes[0].settings.copy()
    return lib.Batch(scenes, **kw)


SCENES = {
    "mixed_events_sleep": lambda: [S.sleeping_pile(3, levels=2, seed=5), S.trigger_scene(60, seed=3)],
    "joints": lambda: [S.jointed_tower(6, seed=4), S.sphere_chains(4), S.slider_rigs(3), S.hinge_rigs(3)],
    "concave": lambda: [S.concave_scene("heightfield", n_bodies=16, shapes=(S.BOX, S.HULL), sleeping=True, seed=22),
                        S.concave_scene("mesh", n_bodies=20, trigger_every=5, seed=21)],
}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("graph", [False, True])
def test_restore_replays_the_same_steps(name, graph):
    dev = _batch(SCENES[name]())
    dev.set_contact_events(True)
    dev.step(DT, 40)
    snap = dev.snapshot()
    assert snap.nbytes > 0
    first = _run(dev, 40, graph, True)
    dev.restore(snap)
    again = _run(dev, 40, graph, True)
    _same(first, again, name)
    # a second restore of the same snapshot replays once more
    dev.restore(snap)
    _same(first, _run(dev, 40, graph, True), name + " (second restore)")
    assert dev.counts()["steps"] == 80
    snap.close()
    dev.close()


def test_changes_after_the_snapshot_are_undone():
    scenes = [S.trigger_scene(40, seed=7), S.jointed_tower(5, seed=8)]
    dev = _batch(scenes)
    dev.step(DT, 20)
    snap = dev.snapshot()
    first = _run(dev, 20, False, False)
    dev.restore(snap)
    # applied changes: every velocity doubled, then stepped; queued changes: calls not yet flushed
    pos, q, v, w = dev.bodies()
    dev.set_bodies(lin=2.0 * v, ang=2.0 * w)
    dev.step(DT, 5)
    dev.body_ops([(lib.OP_SET_LINEAR_VELOCITY, 3, 0, (0.0, 0.0, 40.0), None),
                  (lib.OP_APPLY_FORCE, 5, 0, (0.0, 500.0, 0.0), None)])
    dev.restore(snap)
    _same(first, _run(dev, 20, False, False), "after undone changes")


def test_residual_mode_batch_replays():
    """The fused solver in residual mode re-ranks its world groups every step from the last step's
    iteration counts; that order is part of the restored state."""
    dev = lib.Batch(S.batched_worlds(64, seed=6))
    dev.set_residual_mode(1e-4, 100)
    dev.step(DT, 15)
    snap = dev.snapshot()
    first = []
    for _ in range(15):
        dev.step_graph(DT, 1)
        it, res = dev.world_iterations()
        first.append({"bodies": np.concatenate([a.ravel() for a in dev.bodies()]), "it": it, "res": res})
    dev.restore(snap)
    again = []
    for _ in range(15):
        dev.step_graph(DT, 1)
        it, res = dev.world_iterations()
        again.append({"bodies": np.concatenate([a.ravel() for a in dev.bodies()]), "it": it, "res": res})
    _same(first, again, "residual mode")


def test_snapshot_of_another_batch_is_refused():
    a = lib.Batch(S.box_pile(3, 2, seed=1))
    b = lib.Batch(S.box_pile(3, 2, seed=1))
    snap = a.snapshot()
    with pytest.raises(RuntimeError, match="another batch"):
        b.restore(snap)
    a.close()
    # the snapshot outlives its batch and is still freed cleanly
    with pytest.raises(RuntimeError, match="another batch"):
        b.restore(snap)
    snap.close()


def test_restore_after_an_overflow():
    """A pair table overflow ends a step with an error; restoring the snapshot of the step before gives
    back that state (status clear), and the same step overflows the same way again."""
    dev = lib.Batch(S.box_pile(10, 10, seed=1), pairs_per_collider=1)   # 4096 pairs at most
    snap, state, err = None, None, None
    for k in range(600):
        if snap is not None:
            snap.close()
        snap = dev.snapshot()
        state = np.concatenate([a.ravel() for a in dev.bodies()])
        try:
            dev.step(DT)
        except RuntimeError as e:
            err = (k, str(e))
            break
    assert err is not None and "pair table overflow" in err[1]
    dev.restore(snap)
    assert dev.counts()["status"] == 0
    assert dev.counts()["steps"] == err[0]
    assert np.array_equal(np.concatenate([a.ravel() for a in dev.bodies()]).view(np.uint32), state.view(np.uint32))
    with pytest.raises(RuntimeError, match="pair table overflow"):
        dev.step(DT)
