"""C ABI checks that need no GPU: the library loads, exports every declared symbol, and the host-side
world construction validates its arguments the way the reference's factories do."""
import ctypes as C
import os
import re

import pytest

from reactphysics3d_amd import lib
from reactphysics3d_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rp3d_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(rp3dg_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(lib.EXPORTED) == declared_symbols()


def test_library_exports_every_declared_symbol():
    L = lib.load()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_settings_defaults_match_world_settings():
    # PhysicsWorld::WorldSettings() (include/reactphysics3d/engine/PhysicsWorld.h:122-137)
    L = lib.load()
    s = lib.WorldSettings()
    L.rp3dg_world_settings_default(C.byref(s))
    assert list(s.gravity) == pytest.approx([0.0, -9.81, 0.0])
    assert s.persistent_contact_distance_threshold == pytest.approx(0.03)
    assert s.default_friction_coefficient == pytest.approx(0.3)
    assert s.default_bounciness == pytest.approx(0.5)
    assert s.restitution_velocity_threshold == pytest.approx(0.5)
    assert s.cos_angle_similar_contact_manifold == pytest.approx(0.95)
    assert (s.default_velocity_solver_iterations, s.default_position_solver_iterations) == (6, 3)
    assert s.is_sleeping_enabled == 1


def test_invalid_arguments_are_rejected():
    L = lib.load()
    w = L.rp3dg_world_create(None)
    f3 = (C.c_float * 3)(0, 0, 0)
    q = (C.c_float * 4)(0, 0, 0, 1)
    b = L.rp3dg_body_create(w, f3, q)
    assert b == 0
    assert L.rp3dg_body_set_type(w, 5, 0) < 0            # no such body
    assert L.rp3dg_body_set_type(w, b, 7) < 0            # no such type
    bad = (C.c_float * 4)(-1.0, 0, 0, 0)
    assert L.rp3dg_collider_add(w, b, S.SPHERE, bad, f3, q) < 0
    assert "radius" in lib.last_error()
    assert L.rp3dg_collider_add(w, b, 9, bad, f3, q) == -5  # unsupported shape kind
    ok = (C.c_float * 4)(0.5, 0, 0, 0)
    c = L.rp3dg_collider_add(w, b, S.SPHERE, ok, f3, q)
    assert c == 0
    assert L.rp3dg_collider_set_material(w, c, 0.5, 2.0, 1.0) < 0   # bounciness > 1
    assert L.rp3dg_joint_create(w, 0, b, b, f3, f3, 0, 0, None) < 0  # same body twice
    L.rp3dg_world_destroy(w)


def test_joint_limits_are_validated():
    """SliderJoint.cpp:41-43 asserts min <= 0 <= max and a max motor force >= 0; the C-ABI rejects
    them, at creation and through rp3dg_joint_set_limits, which only acts before the first update."""
    L = lib.load()
    w = L.rp3dg_world_create(None)
    f3 = (C.c_float * 3)(0, 0, 0)
    ax = (C.c_float * 3)(1, 0, 0)
    q = (C.c_float * 4)(0, 0, 0, 1)
    b1, b2 = L.rp3dg_body_create(w, f3, q), L.rp3dg_body_create(w, f3, q)
    above = (C.c_float * 4)(0.2, 1.0, 0.0, 0.0)
    assert L.rp3dg_joint_create(w, S.SLIDER, b1, b2, f3, ax, 0, 1, above) == -1
    assert "min <= 0 <= max" in lib.last_error()
    assert L.rp3dg_joint_create(w, S.SLIDER, b1, b2, f3, ax, 0, 0, above) == 0     # limit off: not checked
    neg_force = (C.c_float * 4)(-1.0, 1.0, 0.5, -2.0)
    assert L.rp3dg_joint_set_limits(w, 0, 2, neg_force) == -1
    ok = (C.c_float * 4)(-1.0, 1.0, 0.5, 2.0)
    assert L.rp3dg_joint_set_limits(w, 0, 3, ok) == 0
    assert L.rp3dg_joint_set_limits(w, 1, 3, ok) == -1                             # no such joint
    cone = (C.c_float * 4)(0.4, 0.0, 0.0, 0.0)
    j = L.rp3dg_joint_create(w, S.BALLSOCKET, b1, b2, f3, f3, 0, 0, None)
    assert j == 1 and L.rp3dg_joint_set_limits(w, j, 1, cone) == 0
    L.rp3dg_world_destroy(w)


def test_hull_rejects_open_mesh():
    L = lib.load()
    w = L.rp3dg_world_create(None)
    verts = (C.c_float * 12)(0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1)
    counts = (C.c_int32 * 3)(3, 3, 3)
    idx = (C.c_int32 * 9)(0, 2, 1, 0, 1, 3, 0, 3, 2)
    assert L.rp3dg_hull_create(w, verts, 4, counts, idx, 3) < 0
    counts4 = (C.c_int32 * 4)(3, 3, 3, 3)
    idx4 = (C.c_int32 * 12)(0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3)
    assert L.rp3dg_hull_create(w, verts, 4, counts4, idx4, 4) == 0
    L.rp3dg_world_destroy(w)


def _prism(n):
    """An n-gon prism: vertices (bottom ring, top ring), two n-vertex caps and n quads, CCW outside."""
    import numpy as np
    a = 2.0 * np.pi * np.arange(n) / n
    ring = np.stack([np.cos(a), np.sin(a)], 1)
    verts = np.concatenate([np.c_[ring, -np.ones(n)], np.c_[ring, np.ones(n)]]).astype(np.float32)
    faces = [list(range(n - 1, -1, -1)), list(range(n, 2 * n))]
    faces += [[i, (i + 1) % n, n + (i + 1) % n, n + i] for i in range(n)]
    counts = np.array([len(f) for f in faces], np.int32)
    idx = np.concatenate([np.array(f, np.int32) for f in faces])
    return verts, counts, idx


def test_hull_face_size_limit():
    """Faces of up to 24 vertices are accepted (the face clipping holds incident face + reference
    face sides = 48 vertices); a 25-vertex cap is refused with RP3DG_EUNSUPPORTED."""
    L = lib.load()
    w = L.rp3dg_world_create(None)
    for n, want in ((24, 0), (25, -5)):
        v, c, i = _prism(n)
        r = L.rp3dg_hull_create(w, lib._fp(v), len(v), lib._ip(c), lib._ip(i), len(c))
        assert (r >= 0) if want == 0 else r == want, (n, r, L.rp3dg_last_error())
    L.rp3dg_world_destroy(w)


def test_scene_loader_runs_on_host():
    sc = S.box_pile(2, 2)
    w = lib.world_from_scene(sc)
    assert w
    lib.load().rp3dg_world_destroy(w)
