"""The narrowphase's degenerate branches and the device's overflow report, against the reference:
coincident sphere centres, parallel capsules (apart and with overlapping inner segments), a capsule
lying on a convex-mesh face (GJK margin and SAT face cases), a cone limit anchored at its body's centre of
mass, two 24-gon prisms stacked cap to cap (the largest clipped polygon the face clipping holds), bit
for bit with the contact and joint order fed to the oracle, and a pair table too small for a settling
pile (the step reports RP3DG_EOVERFLOW with the reason instead of writing past the table)."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests import edge_scenes as E
from tests.test_gpu_joints import _run

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.mark.parametrize("name,make,steps,points", [
    ("coincident_spheres", E.coincident_spheres, 20, 1),
    ("parallel_capsules", lambda: E.parallel_capsules(0.3), 20, 2),
    ("parallel_capsules_coincident", lambda: E.parallel_capsules(0.0, coincident=True), 20, 2),
    ("capsule_on_face_gjk", lambda: E.capsule_on_mesh_face(0.02), 30, 2),
    ("capsule_on_face_sat", lambda: E.capsule_on_mesh_face(0.2), 30, 2),
    ("cone_anchor_at_com", E.cone_anchor_at_com, 120, None),
    ("stacked_24gon_prisms", E.stacked_prisms, 60, 4),
])
@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_degenerate_cases_match_reference(name, make, steps, points, solver):
    from reactphysics3d_amd.lib import Batch
    if points is not None:   # the first step's manifold has the degenerate branch's point count
        dev = Batch(make(), solver=solver)
        dev.step(DT)
        assert dev.contacts()["nb_points"].tolist() == [points], name
        dev.close()
    dev, worst = _run(make(), steps, tol=0.0, solver=solver)
    print(name, solver, "worst", worst)


def test_pair_table_overflow_is_reported():
    from reactphysics3d_amd import lib
    from reactphysics3d_amd.lib import Batch
    dev = Batch(S.box_pile(10, 10, seed=1), pairs_per_collider=1)   # 4096 pairs at most
    err = None
    for k in range(600):
        try:
            dev.step(DT)
        except RuntimeError as e:
            err = (k, str(e))
            break
    assert err is not None, "the settling pile never exceeded 4096 pairs"
    assert "pair table overflow" in err[1]
    c = dev.counts()
    assert c["status"] & 1 and c["nb_pairs"] <= 4096
    for a in dev.bodies():
        assert np.all(np.isfinite(a))
    print("overflow reported at step", err[0], c)
