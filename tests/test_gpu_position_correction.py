"""Position correction techniques other than the defaults, against the reference set the same way:
Baumgarte contacts (PhysicsWorld::setContactsPositionCorrectionTechnique(BAUMGARTE_CONTACTS): the
penetration bias joins the normal impulse and no split velocities are integrated,
ContactSolverSystem.cpp:486-590, DynamicsSystem.cpp:44-71), in both solver layouts, bit for bit with
the device's contact and joint order fed to the oracle; Baumgarte joints
(JointInfo::positionCorrectionTechnique = BAUMGARTE_JOINTS), alone and mixed with non-linear Gauss-Seidel
joints, in one world and over a batch."""
import copy

import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.test_gpu_joints import _run

pytestmark = pytest.mark.gpu


def _baumgarte_contacts(dev, ref):
    dev.set_contacts_baumgarte(True)
    ref.set_contacts_baumgarte(True)


@pytest.mark.parametrize("solver", ["fused", "per_color"])
@pytest.mark.parametrize("name", ["pile", "mixed", "hulls_chains"])
def test_baumgarte_contacts_match_reference(name, solver):
    sc = {"pile": S.box_pile(4, 3, seed=11),
          "mixed": S.mixed_container(300, extent=(12.0, 12.0, 8.0), seed=3),
          "hulls_chains": S.hull_scene(n_hulls=60, n_chains=2, chain_length=5, extent=(8.0, 8.0, 6.0), seed=5)}[name]
    dev, worst = _run(sc, 90, solver=solver, setup=_baumgarte_contacts)
    print(name, solver, "worst", worst)


# ---- Baumgarte joints (JointInfo::positionCorrectionTechnique = BAUMGARTE_JOINTS) ----

def _with_baumgarte(sc, pick=lambda k, joint: True):
    """The scene with RP3DG_JOINT_BAUMGARTE set on the joints that `pick(index, joint)` selects."""
    out = copy.deepcopy(sc)
    out.joints = [j[:6] + (j[6] | S.JOINT_BAUMGARTE,) + j[7:] if pick(k, j) else j for k, j in enumerate(sc.joints)]
    assert any(j[6] & S.JOINT_BAUMGARTE for j in out.joints)
    return out


def _joint_scene(name):
    lim = {S.HINGE: (3, (-0.2, 0.2, 0.5, 2.0)), S.SLIDER: (3, (-0.1, 0.1, -0.2, 5.0)),
           S.BALLSOCKET: (1, (0.15, 0.0, 0.0, 0.0))}
    return {"chains": lambda: S.sphere_chains(n_chains=9, length=8, seed=31),
            "cones": lambda: S.cone_chains(n_chains=6, length=6, seed=7),
            "tower": lambda: S.jointed_tower(height=12, seed=43, kinds=(S.HINGE, S.SLIDER, S.BALLSOCKET, S.FIXED),
                                             limits=lim),
            "sliders": lambda: S.slider_rigs(n_rigs=8, seed=8),
            "hinges": lambda: S.hinge_rigs(n_rigs=8, seed=9)}[name]()


@pytest.mark.parametrize("solver", ["fused", "per_color"])
@pytest.mark.parametrize("name", ["chains", "cones", "tower", "sliders", "hinges"])
def test_baumgarte_joints_match_reference(name, solver):
    """Every joint Baumgarte: bias rows from the step's position error (SolveBallAndSocketJointSystem.cpp:
    113-149, SolveFixedJointSystem.cpp:110-132, SolveHingeJointSystem.cpp:91-209, SolveSliderJointSystem.cpp:
    116-213) and no position solve."""
    dev, worst = _run(_with_baumgarte(_joint_scene(name)), 120, solver=solver)
    print(name, solver, "worst", worst)


@pytest.mark.parametrize("solver", ["fused", "per_color"])
@pytest.mark.parametrize("name", ["tower", "sliders"])
def test_mixed_joint_techniques_match_reference(name, solver):
    """Every other joint Baumgarte: the sliders solved after the first Baumgarte slider of the world get
    no position correction, whatever their own technique (SolveSliderJointSystem.cpp:527 returns)."""
    # sliders: the rigs' first sliders (even creation index) share colour 0 and are solved before the
    # second ones; 2, 6, 10 Baumgarte leaves 0 the only slider corrected in position
    pick = (lambda k, j: k % 4 == 2) if name == "sliders" else (lambda k, j: k % 2 == 1)
    dev, worst = _run(_with_baumgarte(_joint_scene(name), pick), 120, solver=solver)
    print(name, solver, "worst", worst)


@pytest.mark.parametrize("solver", ["fused", "per_color"])
def test_baumgarte_sliders_in_some_worlds_of_a_batch(solver):
    """Worlds of one batch with and without Baumgarte sliders: the slider cut-off is per world."""
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    scenes = [_with_baumgarte(S.slider_rigs(n_rigs=6, seed=21), lambda k, j: k % 3 == 2),
              S.slider_rigs(n_rigs=6, seed=22),
              _with_baumgarte(_joint_scene("tower"), lambda k, j: j[0] == S.SLIDER and k > 4),
              _with_baumgarte(S.hinge_rigs(n_rigs=4, seed=23))]
    for sc in scenes[1:]:   # a batch shares one WorldSettings
        sc.settings = scenes[0].settings.copy()
    dev = Batch(scenes, solver=solver)
    refs = [RefWorld(sc) for sc in scenes]
    order = dev.joint_order()
    for w, ref in enumerate(refs):
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        ref.set_joint_order(order[(order >= j0) & (order < j1)] - j0)
    dt = 1.0 / 60.0
    for k in range(120):
        dev.step(dt)
        ct = dev.contacts()
        pos, q, v, w_ = dev.bodies()
        for wi, ref in enumerate(refs):
            c0, c1 = dev.collider_base[wi], dev.collider_base[wi + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            ref.step_ordered(dt, ct["colliders"][sel, 0] - c0, ct["colliders"][sel, 1] - c0)
            b0, b1 = dev.body_base[wi], dev.body_base[wi + 1]
            dev_state = np.concatenate([a[b0:b1].ravel() for a in (pos, q, v, w_)])
            ref_state = np.concatenate([a.ravel() for a in ref.bodies()])
            err = float(np.max(np.abs(dev_state - ref_state)))
            assert err == 0.0, f"step {k} world {wi}: body state differs by {err}"
