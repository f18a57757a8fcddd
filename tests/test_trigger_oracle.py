"""Trigger colliders in the oracle (the reference run through the harness), pinned by what the
mathematics fixes: a body whose only collider is a trigger falls freely through a static box (the
symplectic-Euler free fall of the reference's integrator, no contact), and its trigger events are
OverlapStart on the first step that starts with its sphere overlapping the box (collision detection
runs on the positions the previous step left), OverlapStay while they overlap and OverlapExit on the
first step after, with no contact events."""
import numpy as np

from reactphysics3d_amd import scene as S

DT = 1.0 / 60.0


def test_trigger_sphere_falls_through_a_box_with_overlap_events(ref_available):
    from oracle.ref import RefWorld
    sc = S.Scene("trigger_pin", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    g = sc.add_body(S.STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, S.BOX, (4.0, 4.0, 0.5))
    s = sc.add_body(S.DYNAMIC, (0.0, 0.0, 3.0))
    sc.add_collider(s, S.SPHERE, (0.5,), trigger=True)
    ref = RefWorld(sc)
    ref.enable_events()
    kinds = []
    z_before = 3.0
    for k in range(1, 80):
        ref.step(DT)
        p, _, v, _ = ref.bodies()
        assert abs(v[1, 2] - (-9.81 * k * DT)) <= 1e-4 * k   # free fall: the trigger pushes nothing
        z = 3.0 - 9.81 * DT * DT * k * (k + 1) / 2
        assert abs(p[1, 2] - z) <= 1e-3
        ev = ref.events()
        assert np.all(ev[:, 2] >= 3), "a trigger pair must not report contact events"
        # the step's narrowphase sees the position the previous step left: sphere centre within a radius of
        # the box's z extent [-1, 0]
        overlapping = -1.0 - 0.5 < z_before < 0.0 + 0.5
        kinds.append((overlapping, tuple(ev[:, 2])))
        z_before = z
    first = next(i for i, (o, _) in enumerate(kinds) if o)
    last = max(i for i, (o, _) in enumerate(kinds) if o)
    assert kinds[first][1] == (3,)                       # OverlapStart
    assert all(kinds[i][1] == (4,) for i in range(first + 1, last + 1))   # OverlapStay
    assert kinds[last + 1][1] == (5,)                    # OverlapExit
    assert all(kinds[i][1] == () for i in range(first)) and all(kinds[i][1] == () for i in range(last + 2, len(kinds)))
