"""Device joint solvers (ball-and-socket with cone limits, fixed, hinge and slider with limits and
motors; velocity + non-linear Gauss-Seidel position correction) vs the reference fed the device's joint
and contact order."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts, pair_set

pytestmark = pytest.mark.gpu


# The contract for body state is 1e-4; the device joint rows keep the reference's operation order and
# its C library's acos / atan2 bits, so these scenes hold the stronger bit-for-bit equality.
def _run(sc, steps, tol=0.0, solver="auto", setup=None):
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    dev = Batch(sc, solver=solver)
    ref = RefWorld(sc)
    ref.set_joint_order(dev.joint_order())
    if setup:
        setup(dev, ref)
    dt = 1.0 / 60.0
    worst = 0.0
    for k in range(steps):
        dev.step(dt)
        ct = dev.contacts()
        ref.step_ordered(dt, ct["colliders"][:, 0], ct["colliders"][:, 1])
        d1, d2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert pair_set(d1, d2) == pair_set(r1, r2), f"step {k}: oriented pair sets differ"
        errs = compare_contacts(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        dev_state = np.concatenate([a.ravel() for a in dev.bodies()])
        ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
        worst = max(worst, float(np.max(np.abs(dev_state - ref_state))))
        assert worst <= tol, f"step {k}: body state differs by {worst}"
    return dev, worst


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_ball_socket_chains(solver):
    sc = S.sphere_chains(n_chains=9, length=8, seed=31)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_hinge_fixed_tower(solver):
    sc = S.jointed_tower(height=8, seed=41)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_slider_rigs_limits_motors(solver):
    sc = S.slider_rigs(n_rigs=8, seed=8)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_hinge_rigs_limits_motors(solver):
    sc = S.hinge_rigs(n_rigs=8, seed=9)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_ball_socket_cone_limits(solver):
    sc = S.cone_chains(n_chains=6, length=6, seed=7)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_tower_of_all_joint_kinds(solver):
    # every joint kind in one colour-sorted table, with contacts between the unjointed boxes
    lim = {S.HINGE: (3, (-0.2, 0.2, 0.5, 2.0)), S.SLIDER: (3, (-0.1, 0.1, -0.2, 5.0)),
           S.BALLSOCKET: (1, (0.15, 0.0, 0.0, 0.0))}
    sc = S.jointed_tower(height=12, seed=43, kinds=(S.HINGE, S.SLIDER, S.BALLSOCKET, S.FIXED), limits=lim)
    dev, worst = _run(sc, 120, solver=solver)
    print("worst", worst, dev.counts())
