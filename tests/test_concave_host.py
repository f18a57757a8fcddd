"""The middle phase of concave shapes on the host, against the reference (CPU): for seeded AABBs in a
concave collider's local space, rp3dg_concave_overlapping_triangles lists the same triangles in the same
order, with bit-identical vertices and vertex normals, as ConcaveShape::computeOverlappingTriangles of the
compiled reference (ConcaveMeshShape.cpp:148-176 over the mesh's DynamicAABBTree; HeightField.cpp:133-180).
That pins the tree the library builds (TriangleMesh::initBVHTree, DynamicAABBTree insertion and
rotations), the vertex clean-up and vertex normals (TriangleMesh.cpp:100-318), the scaled normals
(ConcaveMeshShape.cpp:53-70) and the height field grid walk."""
import ctypes as C

import numpy as np
import pytest

from reactphysics3d_amd import scene as S

pytestmark = pytest.mark.usefixtures("ref_available")


def _host_overlap(L, w, collider, lo, hi, cap=4096):
    ids = np.zeros(cap, np.int32)
    vs = np.zeros((cap, 3, 3), np.float32)
    ns = np.zeros((cap, 3, 3), np.float32)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    n = L.rp3dg_concave_overlapping_triangles(w, collider, fp(lo), fp(hi), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                              fp(vs), fp(ns), cap)
    assert 0 <= n <= cap
    return ids[:n], vs[:n], ns[:n]


def _check(sc, nb_triangles=None, trials=60, seed=3):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import load, world_from_scene
    L = load()
    w = world_from_scene(sc)
    ref = RefWorld(sc)
    rng = np.random.default_rng(seed)
    cols = [i for i, c in enumerate(sc.colliders) if c[1] in (S.CONCAVE_MESH, S.HEIGHTFIELD)]
    total = 0
    try:
        for c in cols:
            kind = sc.colliders[c][1]
            scale = np.array(sc.colliders[c][2][1:4], np.float32)
            for _ in range(trials):
                centre = rng.uniform(-6.0, 6.0, 3) * scale
                half = rng.uniform(0.05, 2.5, 3) * scale
                lo, hi = (centre - half).astype(np.float32), (centre + half).astype(np.float32)
                hid, hv, hn = _host_overlap(L, w, c, lo, hi)
                rid, rv, rn = ref.concave_overlapping_triangles(c, lo, hi)
                if kind == S.CONCAVE_MESH:
                    rid = rid - nb_triangles   # ConcaveMeshShape::computeTriangleShapeId = nb triangles + index
                assert np.array_equal(hid, rid), f"collider {c}: triangle order differs\n{hid}\nvs\n{rid}"
                assert np.array_equal(hv.view(np.uint32), rv.view(np.uint32)), f"collider {c}: vertices differ"
                assert np.array_equal(hn.view(np.uint32), rn.view(np.uint32)), f"collider {c}: normals differ"
                total += len(hid)
    finally:
        L.rp3dg_world_destroy(w)
    assert total > 0
    return total


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (1.7, 0.6, 1.3)])
@pytest.mark.parametrize("normals", [False, True])
def test_triangle_mesh_overlap_order_matches_reference(scale, normals):
    sc = S.concave_scene("mesh", n_bodies=0, nx=14, scale=scale, normals=normals)
    print("triangles", _check(sc, nb_triangles=2 * 14 * 14))


def test_degenerate_triangles_and_unused_vertices_are_cleaned_like_the_reference():
    """Zero-area and zero-edge triangles are dropped and unused vertices removed before the tree is built
    (vertex 0 is kept even when unused, as the reference's loop stops at 1)."""
    v, t = S.terrain_mesh(8, 8, seed=4)
    extra = np.array([[50.0, 50.0, 0.0], [1.0, 1.0, 0.3], [1.0, 1.0, 0.3]], np.float32)
    v = np.concatenate([v, extra])
    n0 = len(v) - 3
    bad = np.array([[0, 1, 1], [n0 + 1, n0 + 2, 5], [2, 3, 4]], np.int32)   # zero edge, zero edge, flat
    # remap so that vertex 0 is unused: shift every index of the good triangles by one where it hits 0
    good = t.copy()
    good[good == 0] = 1
    sc = S.Scene("degenerate_mesh", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    g = sc.add_body(S.STATIC, (0.0, 0.0, 0.0))
    m = sc.add_triangle_mesh(v, np.concatenate([bad, good]))
    sc.add_collider(g, S.CONCAVE_MESH, (float(m), 1.0, 1.0, 1.0))
    from reactphysics3d_amd.lib import load, world_from_scene
    L = load()
    w = world_from_scene(sc)
    L.rp3dg_world_destroy(w)
    # triangles that survive: the valid ones among bad + good (good ones whose two corners became 1 drop)
    vv = v.astype(np.float32)
    tris = np.concatenate([bad, good])
    eps = 3 * np.abs(vv).max(axis=0).sum(dtype=np.float32) * np.float32(np.finfo(np.float32).eps)
    keep = 0
    for a, b, c in tris:
        e = [vv[b] - vv[a], vv[c] - vv[b], vv[a] - vv[c]]
        area = np.cross(vv[c] - vv[a], vv[b] - vv[a])
        if (area @ area) >= eps * eps and all((x @ x) >= eps * eps for x in e):
            keep += 1
    print("valid triangles", keep, "of", len(tris))
    _check(sc, nb_triangles=keep)


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (1.5, 0.8, 1.2)])
def test_heightfield_overlap_order_matches_reference(scale):
    sc = S.concave_scene("heightfield", n_bodies=0, nx=13, scale=scale)
    print("triangles", _check(sc))
