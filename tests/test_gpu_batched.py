"""Many uneven worlds in one batch (box piles and stacks with hinge / fixed joints): every world is
checked against its own reference PhysicsWorld fed that world's share of the device's contact and
joint order, for both solver layouts; graph replay is checked against eager stepping."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _world_views(dev, ct, order):
    """Splits global contacts / joint order into per-world, world-local pieces (order preserved)."""
    out = []
    for w in range(len(dev.scenes)):
        c0, c1 = dev.collider_base[w], dev.collider_base[w + 1]
        sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
        local = {k: a[sel] for k, a in ct.items()}
        local["colliders"] = local["colliders"] - c0
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        jo = order[(order >= j0) & (order < j1)] - j0
        out.append((local, jo))
    return out


def _run(scenes, steps, solver, tol=1e-4, group_bodies=0):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(scenes, solver=solver, group_bodies=group_bodies)
    refs = [RefWorld(sc) for sc in scenes]
    order = dev.joint_order()
    for (_, jo), ref in zip(_world_views(dev, {"colliders": np.zeros((0, 2), np.int32)}, order), refs):
        ref.set_joint_order(jo)
    worst = 0.0
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        views = _world_views(dev, ct, order)
        p1, p2 = dev.pairs()
        pos, q, v, w = dev.bodies()
        for wi, (ref, (local, _)) in enumerate(zip(refs, views)):
            ref.step_ordered(DT, local["colliders"][:, 0], local["colliders"][:, 1])
            c0, c1 = dev.collider_base[wi], dev.collider_base[wi + 1]
            sel = (p1 >= c0) & (p1 < c1)
            r1, r2 = ref.pairs()
            assert pair_set(p1[sel] - c0, p2[sel] - c0) == pair_set(r1, r2), f"step {k} world {wi}: pair sets differ"
            errs = compare_contacts(local, ref.contacts())
            assert not errs, f"step {k} world {wi}: " + "; ".join(errs[:5])
            b0, b1 = dev.body_base[wi], dev.body_base[wi + 1]
            dev_state = np.concatenate([a[b0:b1].ravel() for a in (pos, q, v, w)])
            ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
            worst = max(worst, float(np.max(np.abs(dev_state - ref_state))))
            assert worst <= tol, f"step {k} world {wi}: body state differs by {worst}"
    return dev, worst


@pytest.mark.parametrize("solver,group_bodies", [("fused", 0), ("fused", 1), ("per_color", 0)])
def test_uneven_worlds_match_reference(solver, group_bodies):
    # fused, 0: worlds packed into groups of up to 256 bodies; fused, 1: one world per workgroup
    scenes = [S.batched_world(w, seed=100, min_bodies=8, max_bodies=48, max_joints=8) for w in range(12)]
    dev, worst = _run(scenes, 60, solver, group_bodies=group_bodies)
    c = dev.counts()
    print(solver, "worst", worst, c)
    assert c["nb_worlds"] == 12 and c["nb_manifolds"] > 0 and c["nb_joints"] > 0


def test_graph_replay_matches_eager():
    from reactphysics3d_amd.lib import Batch
    scenes = [S.batched_world(w, seed=7, min_bodies=8, max_bodies=96, max_joints=16) for w in range(24)]
    eager = Batch(scenes, solver="fused")
    graph = Batch(scenes, solver="fused")
    for _ in range(5):
        eager.step(DT)
        graph.step(DT)
    for rep in range(6):
        eager.step(DT, 2)
        graph.step_graph(DT, 2)
        for a, b in zip(eager.bodies(), graph.bodies()):
            assert np.array_equal(a, b), f"replay {rep}: graph state differs from eager"
        assert eager.counts() == graph.counts()
        # eager work between replays on the graph batch itself
        graph.step(DT)
        eager.step(DT)


def test_residual_mode_converges():
    from reactphysics3d_amd.lib import Batch
    scenes = [S.batched_world(w, seed=11, min_bodies=16, max_bodies=128, max_joints=8) for w in range(16)]
    dev = Batch(scenes, solver="fused")
    dev.step(DT, 30)
    dev.set_residual_mode(1e-4, 200)
    dev.step(DT)
    st = dev.iteration_stats()
    print(st)
    assert 1 <= st["mean_iterations"] <= st["max_iterations"] <= 200
    # every world either reached the threshold or ran the cap
    if st["max_iterations"] < 200:
        assert st["max_residual"] < 1e-4
    dev.set_residual_mode(0.0)
    dev.step(DT)
    assert dev.iteration_stats()["max_iterations"] == 10


@pytest.mark.parametrize("residual", [0.0, 1e-4])
def test_world_groups_do_not_change_results(residual):
    """Packing worlds into one workgroup changes nothing: per-world results, iteration counts and
    residuals are identical to one world per workgroup, in fixed and residual mode (a converged world
    of a group skips the sweeps its neighbours still run), and so does the order the workgroups are
    dispatched in (by the previous step's sweep counts, or by size)."""
    from reactphysics3d_amd.lib import Batch
    scenes = [S.batched_world(w, seed=21, min_bodies=8, max_bodies=160, max_joints=8) for w in range(40)]
    runs = []
    for gb, by_sweeps in ((1, True), (128, True), (512, True), (128, False)):
        dev = Batch(scenes, solver="fused", group_bodies=gb)
        dev.set_group_order(by_sweeps)   # workgroup dispatch order: longest expected first, or static
        if residual > 0:
            dev.set_residual_mode(residual, 150)
        stats = []
        for _ in range(4):
            dev.step(DT, 5)
            stats.append(dev.iteration_stats())
        runs.append((dev.bodies(), stats, dev.counts()))
    base = runs[0]
    for bodies, stats, counts in runs[1:]:
        for a, b in zip(base[0], bodies):
            assert np.array_equal(a, b), "grouped worlds differ from single-world workgroups"
        assert stats == base[1]
        assert counts == base[2]
    if residual > 0:
        assert base[1][-1]["max_iterations"] < 150 or base[1][-1]["max_residual"] >= residual
        assert base[1][-1]["mean_iterations"] < base[1][-1]["max_iterations"]


@pytest.mark.parametrize("kind", ["batched", "mixed"])
def test_world_scan_broadphase_matches_tree(kind):
    """The per-world box scan finds the pair set of the tree query: identical pair tables, contacts and
    body states after every step (dense piles make some colliders exceed the eight-partner slab)."""
    from reactphysics3d_amd.lib import Batch
    if kind == "batched":
        scenes = [S.batched_world(w, seed=31, min_bodies=8, max_bodies=300, max_joints=8) for w in range(24)]
    else:
        scenes = [S.mixed_container(400, extent=(8.0, 8.0, 10.0), seed=5)]
    tree = Batch(scenes, solver="fused", broadphase="tree")
    scan = Batch(scenes, solver="fused", broadphase="world_scan")
    for step in range(40):
        tree.step(DT)
        scan.step(DT)
        pt, ps = tree.pairs(), scan.pairs()
        assert all(np.array_equal(a, b) for a, b in zip(pt, ps)), f"step {step}: pair tables differ"
        for a, b in zip(tree.bodies(), scan.bodies()):
            assert np.array_equal(a, b), f"step {step}: body states differ"
    assert tree.counts() == scan.counts()
