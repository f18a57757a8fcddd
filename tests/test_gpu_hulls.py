"""Convex-mesh bodies (hulls of 24 ellipsoid points, polygon input shared with the reference) and
sphere chains: hull-vs-hull SAT, hull-vs-box, sphere-vs-hull, ball-and-socket joints, vs the
reference fed the device's contact and joint order."""
import pytest

from reactphysics3d_amd import scene as S
from tests.test_gpu_joints import _run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,solver", [(5, "fused"), (8, "per_color")])
def test_hull_scene_matches_reference(seed, solver):
    sc = S.hull_scene(n_hulls=120, n_chains=3, chain_length=5, extent=(10.0, 10.0, 8.0), seed=seed)
    dev, worst = _run(sc, 90, solver=solver)
    c = dev.counts()
    print("worst", worst, c)
    assert c["nb_tests"][6] > 0
