"""Pins the oracle (the compiled reference + harness) against the reference's own narrowphase
known-answer tests: each case, stepped once without gravity, reports the expected contact points."""
import pytest

from tests import kat


@pytest.mark.parametrize("case", kat.load()["cases"], ids=lambda c: c["name"])
def test_oracle_matches_reference_kat(case, ref_available):
    from oracle.ref import RefWorld
    sc, _ = kat.scene(case)
    ref = RefWorld(sc)
    ref.step(1.0 / 60.0)
    problems = kat.check(case, ref.contacts(), kat.load()["epsilon"])
    assert not problems, problems
