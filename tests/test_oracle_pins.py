"""Pins the rest of the oracle with checks the mathematics fixes: symplectic-Euler free fall,
box / sphere mass properties, and the coloured-order harness reducing to the reference's own step
when it is fed the reference's own solve order."""
import numpy as np
import pytest

from reactphysics3d_amd import scene as S

DT = 1.0 / 60.0


def test_free_fall_closed_form(ref_available):
    from oracle.ref import RefWorld
    sc = S.Scene("fall", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 100.0))
    sc.add_collider(b, S.SPHERE, (0.5,))
    ref = RefWorld(sc)
    for k in range(1, 61):
        ref.step(DT)
        p, _, v, _ = ref.bodies()
        # v_k = -g k dt, x_k = x_0 - g dt^2 k (k + 1) / 2 (velocity first, then position)
        assert abs(v[0, 2] - (-9.81 * k * DT)) <= 1e-4 * k
        assert abs(p[0, 2] - (100.0 - 9.81 * DT * DT * k * (k + 1) / 2)) <= 1e-3


def test_mass_properties_closed_form(ref_available):
    from oracle.ref import RefWorld
    sc = S.Scene("mass", S.make_settings())
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 0.0))
    sc.add_collider(b, S.BOX, (0.5, 1.0, 1.5), density=2.0)
    s = sc.add_body(S.DYNAMIC, (5.0, 0.0, 0.0))
    sc.add_collider(s, S.SPHERE, (0.7,), density=3.0)
    m, inertia, com = RefWorld(sc).mass_properties()
    mb = 2.0 * 8 * 0.5 * 1.0 * 1.5
    assert m[0] == pytest.approx(mb, rel=1e-6)
    # box: I = m/3 (b^2 + c^2) about each axis with half-extents (a, b, c)
    assert inertia[0] == pytest.approx([mb / 3 * (1.0 + 2.25), mb / 3 * (0.25 + 2.25), mb / 3 * (0.25 + 1.0)], rel=1e-5)
    ms = 3.0 * 4.0 / 3.0 * np.pi * 0.7 ** 3
    assert m[1] == pytest.approx(ms, rel=1e-5)
    assert inertia[1] == pytest.approx([0.4 * ms * 0.49] * 3, rel=1e-5)


def test_ordered_step_with_reference_order_is_the_reference_step(ref_available):
    from oracle.ref import RefWorld
    sc = S.box_pile(3, 2, seed=21)
    a, b = RefWorld(sc), RefWorld(sc)
    for k in range(40):
        a.step(DT)
        ct = a.contacts()   # manifolds in the order the reference solved them this step
        b.step_ordered(DT, ct["colliders"][:, 0], ct["colliders"][:, 1])
        for x, y in zip(a.bodies(), b.bodies()):
            assert np.array_equal(x, y), f"step {k}"
