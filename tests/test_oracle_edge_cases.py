"""The degenerate narrowphase branches in the oracle, pinned by the geometry: coincident sphere centres
take the normal (0, 1, 0) and the depth r1 + r2; parallel capsules give two points at the ends of the
overlap of their inner segments, at depth r1 + r2 - distance between the axes; a capsule lying on a box
face gives two points, in the GJK margin case and in the SAT face case."""
import numpy as np
import pytest

from tests import edge_scenes as E

DT = 1.0 / 60.0


def _contacts(sc):
    from oracle.ref import RefWorld
    ref = RefWorld(sc)
    ref.step(DT)
    return ref.contacts()


def test_coincident_sphere_centres(ref_available):
    ct = _contacts(E.coincident_spheres())
    assert ct["nb_points"].tolist() == [1]
    n = ct["normal_depth"][0, 0]
    assert np.allclose(n[:3], [0.0, 1.0, 0.0]) or np.allclose(n[:3], [0.0, -1.0, 0.0])
    assert n[3] == pytest.approx(0.5 + 0.8, rel=1e-6)


@pytest.mark.parametrize("gap,coincident", [(0.3, False), (0.0, True)])
def test_parallel_capsules_two_points(ref_available, gap, coincident):
    ct = _contacts(E.parallel_capsules(gap=gap, coincident=coincident))
    assert ct["nb_points"].tolist() == [2]
    depths = ct["normal_depth"][0, :2, 3]
    assert np.allclose(depths, 0.5 + 0.4 - gap, atol=1e-5)
    # the two points sit at the ends of the shared length of the segments: y = offset - 1 and y = 1 in
    # the first capsule's frame (segments [-1, 1] and [offset - 1, offset + 1])
    ys = sorted(ct["local1"][0, :2, 1]) if int(ct["colliders"][0, 0]) == 0 else sorted(ct["local2"][0, :2, 1])
    assert ys == pytest.approx([0.6 - 1.0, 1.0], abs=1e-5)


@pytest.mark.parametrize("depth", [0.02, 0.2])
def test_capsule_on_a_mesh_face_two_points(ref_available, depth):
    ct = _contacts(E.capsule_on_mesh_face(depth))
    assert ct["nb_points"].tolist() == [2]
    assert np.allclose(ct["normal_depth"][0, :2, 3], depth, atol=1e-5)
    assert np.allclose(np.abs(ct["normal_depth"][0, :2, 2]), 1.0, atol=1e-6)   # the face normal, +-z


def test_stacked_24gon_prisms_face_contact(ref_available):
    """Two 24-gon prisms cap to cap, the top one turned by 0.07 rad and sunk 0.005 into the base: the
    face contact clips a 24-vertex cap by 24 side planes (a 48-vertex polygon) and reduces it to 4 points
    on the rim of the overlap, at depth 0.005 along the common axis."""
    ct = _contacts(E.stacked_prisms())
    assert ct["nb_points"].tolist() == [4]
    nd = ct["normal_depth"][0]
    assert np.allclose(np.abs(nd[:, 2]), 1.0, atol=1e-6)
    assert np.allclose(nd[:, 3], 0.005, atol=2e-5)
    # the base is collider 0 at the origin: its local points are on its top cap (z = 0.5), near the rim
    local = ct["local1"][0] if int(ct["colliders"][0, 0]) == 0 else ct["local2"][0]
    r = np.hypot(local[:, 0], local[:, 1])
    assert np.allclose(local[:, 2], 0.5, atol=1e-5)
    assert np.all(r <= 1.0 + 1e-5) and np.all(r >= 0.95)
