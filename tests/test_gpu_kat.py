"""The device narrowphase against the reference's own known-answer contact points."""
import pytest

from tests import kat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", kat.load()["cases"], ids=lambda c: c["name"])
def test_device_matches_reference_kat(case):
    from reactphysics3d_amd.lib import Batch
    sc, _ = kat.scene(case)
    dev = Batch(sc)
    dev.step(1.0 / 60.0)
    problems = kat.check(case, dev.contacts(), kat.load()["epsilon"])
    assert not problems, problems
