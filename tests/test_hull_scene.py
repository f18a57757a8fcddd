"""Convex-mesh scene inputs (host side, no GPU): the hull builder against scipy's Qhull, and the
reference harness accepting the polygons."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S


def test_convex_hull_matches_qhull():
    spatial = pytest.importorskip("scipy.spatial")
    rng = S.SplitMix64(17)
    for _ in range(5):
        v, f = S.ellipsoid_hull(rng)
        h = spatial.ConvexHull(v.astype(np.float64))
        assert len(f) == len(h.simplices)
        # outward orientation: every face normal points away from the centroid; volume by divergence
        c = v.mean(0)
        vol = 0.0
        for a, b, d in f:
            n = np.cross(v[b] - v[a], v[d] - v[a])
            assert np.dot(n, v[a] - c) > 0
            vol += np.dot(v[a], np.cross(v[b], v[d])) / 6.0
        assert abs(vol - h.volume) < 1e-4 * h.volume


def test_reference_accepts_hull_scene(ref_available):
    from oracle.ref import RefWorld
    sc = S.hull_scene(n_hulls=20, n_chains=1, chain_length=3, extent=(8.0, 8.0, 6.0), seed=5)
    r = RefWorld(sc)
    for _ in range(10):
        r.step(1.0 / 60.0)
    p, q, v, w = r.bodies()
    assert np.isfinite(p).all() and np.isfinite(v).all()
