"""Position correction techniques other than the defaults, against the reference set the same way:
Baumgarte contacts (PhysicsWorld::setContactsPositionCorrectionTechnique(BAUMGARTE_CONTACTS): the
penetration bias joins the normal impulse and no split velocities are integrated,
ContactSolverSystem.cpp:486-590, DynamicsSystem.cpp:44-71), in both solver layouts, bit for bit with
the device's contact and joint order fed to the oracle."""
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
