"""ctypes binding of the C ABI (include/rp3d_mi355x.h).

The HIP library must be built (``python -m reactphysics3d_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback for the step path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scene as S

# RP3DG_LIBRARY selects another build of the library (the AddressSanitizer build of the host code)
LIB_PATH = os.environ.get("RP3DG_LIBRARY") or \
    os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "librp3d_mi355x.so")

_lib = None


class WorldSettings(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3),
        ("persistent_contact_distance_threshold", C.c_float),
        ("default_friction_coefficient", C.c_float),
        ("default_bounciness", C.c_float),
        ("restitution_velocity_threshold", C.c_float),
        ("cos_angle_similar_contact_manifold", C.c_float),
        ("default_velocity_solver_iterations", C.c_uint32),
        ("default_position_solver_iterations", C.c_uint32),
        ("is_sleeping_enabled", C.c_int32),
        ("is_gravity_enabled", C.c_int32),
        ("time_before_sleep", C.c_float),
        ("sleep_linear_velocity", C.c_float),
        ("sleep_angular_velocity", C.c_float),
    ]


class Counts(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("nb_bodies", "nb_colliders", "nb_joints", "nb_worlds", "nb_pairs",
                                         "nb_manifolds", "nb_contact_points", "nb_colors", "status")] + \
               [("nb_tests", C.c_int64 * 7), ("total_contact_points", C.c_int64), ("steps", C.c_int64),
                ("nb_new_points", C.c_int64), ("nb_box_tests", C.c_int64)]


class BatchOptions(C.Structure):
    _fields_ = [("solver", C.c_int32), ("fused_block", C.c_int32), ("pairs_per_collider", C.c_int32),
                ("group_bodies", C.c_int32), ("broadphase", C.c_int32), ("reserved", C.c_int32 * 3)]


SOLVER_MODES = {"auto": 0, "per_color": 1, "fused": 2}


class ContactView(C.Structure):
    _fields_ = [("colliders", C.POINTER(C.c_int32)), ("nb_points", C.POINTER(C.c_int32)),
                ("color", C.POINTER(C.c_int32)), ("normal_depth", C.POINTER(C.c_float)),
                ("local1", C.POINTER(C.c_float)), ("local2", C.POINTER(C.c_float)),
                ("impulse", C.POINTER(C.c_float)), ("friction", C.POINTER(C.c_float)),
                ("manifold", C.POINTER(C.c_int32))]


# Every symbol the header declares (the not-GPU test checks that each is exported).
EXPORTED = [
    "rp3dg_world_settings_default", "rp3dg_world_create", "rp3dg_world_destroy", "rp3dg_world_set_iterations",
    "rp3dg_world_set_gravity", "rp3dg_body_create", "rp3dg_body_set_type", "rp3dg_body_set_velocity",
    "rp3dg_body_set_damping", "rp3dg_body_enable_gravity", "rp3dg_body_apply_force_torque",
    "rp3dg_body_set_lock_factors", "rp3dg_body_update_mass_properties", "rp3dg_body_set_mass_properties",
    "rp3dg_collider_add", "rp3dg_collider_set_material", "rp3dg_collider_set_filter", "rp3dg_hull_create",
    "rp3dg_triangle_mesh_create", "rp3dg_heightfield_create", "rp3dg_concave_overlapping_triangles",
    "rp3dg_joint_create", "rp3dg_joint_set_limits", "rp3dg_body_set_allowed_to_sleep", "rp3dg_batch_get_sleeping",
    "rp3dg_world_get_body_sleeping", "rp3dg_batch_set_contact_events", "rp3dg_batch_get_contact_events",
    "rp3dg_world_set_contact_events", "rp3dg_world_get_batch", "rp3dg_batch_raycast", "rp3dg_batch_apply_forces",
    "rp3dg_world_update", "rp3dg_batch_create", "rp3dg_batch_create_ex",
    "rp3dg_batch_destroy", "rp3dg_batch_step",
    "rp3dg_batch_step_graph", "rp3dg_batch_synchronize", "rp3dg_batch_get_counts", "rp3dg_batch_get_bodies",
    "rp3dg_batch_set_bodies", "rp3dg_batch_get_pairs", "rp3dg_batch_get_contacts", "rp3dg_batch_get_joint_order",
    "rp3dg_batch_profile", "rp3dg_batch_get_kernel_times", "rp3dg_batch_set_residual_mode",
    "rp3dg_batch_get_iteration_stats", "rp3dg_last_error", "rp3dg_version", "rp3dg_mlp_create", "rp3dg_mlp_destroy",
    "rp3dg_mlp_set_params", "rp3dg_mlp_get_params", "rp3dg_mlp_set_normalization", "rp3dg_mlp_forward",
    "rp3dg_mlp_gradient", "rp3dg_mlp_train", "rp3dg_batch_set_warm_start", "rp3dg_batch_set_logging",
    "rp3dg_batch_get_log", "rp3dg_batch_autotune", "rp3dg_batch_get_tune_table", "rp3dg_world_get_body_state",
    "rp3dg_batch_get_world_iterations", "rp3dg_batch_get_warm_start_predictions", "rp3dg_mlp_gradient_noalloc",
    "rp3dg_hull_create_from_points", "rp3dg_quickhull", "rp3dg_collider_set_is_trigger",
    "rp3dg_collider_set_is_simulation_collider", "rp3dg_collider_set_is_world_query_collider",
    "rp3dg_collider_get_kind", "rp3dg_batch_set_solver_shape",
    "rp3dg_batch_body_set_transform", "rp3dg_batch_body_set_velocity", "rp3dg_batch_body_apply_force",
    "rp3dg_batch_body_apply_torque", "rp3dg_batch_body_reset_force_torque", "rp3dg_batch_body_set_sleeping",
    "rp3dg_batch_get_force_torque", "rp3dg_body_set_transform", "rp3dg_body_apply_force", "rp3dg_body_apply_torque",
    "rp3dg_body_reset_force_torque", "rp3dg_body_get_force_torque", "rp3dg_body_set_sleeping",
    "rp3dg_batch_raycast_all", "rp3dg_batch_test_overlap_bodies", "rp3dg_batch_test_overlap",
    "rp3dg_batch_test_collision", "rp3dg_batch_get_query_results", "rp3dg_world_upload",
    "rp3dg_batch_set_group_order", "rp3dg_world_set_contacts_position_correction",
    "rp3dg_batch_set_contacts_position_correction", "rp3dg_batch_snapshot", "rp3dg_batch_restore",
    "rp3dg_snapshot_bytes", "rp3dg_snapshot_destroy",
]

# RigidBody call codes of Batch.body_ops (the order of the C-ABI's per-body calls)
OP_SET_TRANSFORM, OP_SET_LINEAR_VELOCITY, OP_SET_ANGULAR_VELOCITY = 1, 2, 3
OP_APPLY_FORCE, OP_APPLY_TORQUE, OP_RESET_FORCE, OP_RESET_TORQUE, OP_SET_SLEEPING = 4, 5, 6, 7, 8
FORCE_LOCAL, FORCE_AT_POINT, FORCE_POINT_LOCAL = 1, 2, 4

MLP_PARAMS = 20099
N_FEATURES = 24


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP step library not built: {LIB_PATH} missing "
                           "(run `python -m reactphysics3d_amd.build`)")
    L = C.CDLL(LIB_PATH)
    vp, f, i32, i64, u32 = C.c_void_p, C.c_float, C.c_int32, C.c_int64, C.c_uint32
    fptr, iptr = C.POINTER(C.c_float), C.POINTER(C.c_int32)
    sig = {
        "rp3dg_world_settings_default": (None, [C.POINTER(WorldSettings)]),
        "rp3dg_world_create": (vp, [C.POINTER(WorldSettings)]),
        "rp3dg_world_destroy": (None, [vp]),
        "rp3dg_world_set_iterations": (C.c_int, [vp, u32, u32]),
        "rp3dg_world_set_gravity": (C.c_int, [vp, fptr]),
        "rp3dg_body_create": (i32, [vp, fptr, fptr]),
        "rp3dg_body_set_type": (C.c_int, [vp, i32, i32]),
        "rp3dg_body_set_velocity": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_body_set_damping": (C.c_int, [vp, i32, f, f]),
        "rp3dg_body_enable_gravity": (C.c_int, [vp, i32, i32]),
        "rp3dg_body_apply_force_torque": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_body_set_lock_factors": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_body_update_mass_properties": (C.c_int, [vp, i32]),
        "rp3dg_body_set_mass_properties": (C.c_int, [vp, i32, f, fptr, fptr]),
        "rp3dg_collider_add": (i32, [vp, i32, i32, fptr, fptr, fptr]),
        "rp3dg_collider_set_material": (C.c_int, [vp, i32, f, f, f]),
        "rp3dg_collider_set_filter": (C.c_int, [vp, i32, C.c_uint16, C.c_uint16]),
        "rp3dg_hull_create": (i32, [vp, fptr, i32, iptr, iptr, i32]),
        "rp3dg_triangle_mesh_create": (i32, [vp, fptr, i32, iptr, i32, fptr, iptr]),
        "rp3dg_heightfield_create": (i32, [vp, i32, i32, fptr]),
        "rp3dg_concave_overlapping_triangles": (i32, [vp, i32, fptr, fptr, iptr, fptr, fptr, i32]),
        "rp3dg_joint_create": (i32, [vp, i32, i32, i32, fptr, fptr, i32, i32, fptr]),
        "rp3dg_joint_set_limits": (C.c_int, [vp, i32, i32, fptr]),
        "rp3dg_body_set_allowed_to_sleep": (C.c_int, [vp, i32, i32]),
        "rp3dg_batch_get_sleeping": (C.c_int, [vp, iptr, C.c_int64]),
        "rp3dg_world_get_body_sleeping": (C.c_int, [vp, i32]),
        "rp3dg_batch_set_contact_events": (C.c_int, [vp, i32]),
        "rp3dg_world_set_contact_events": (C.c_int, [vp, i32]),
        "rp3dg_world_get_batch": (vp, [vp]),
        "rp3dg_batch_raycast": (C.c_int, [vp, C.c_int64, iptr, fptr, fptr, fptr, C.c_uint16, iptr, fptr, fptr, fptr]),
        "rp3dg_batch_apply_forces": (C.c_int, [vp, fptr, fptr]),
        "rp3dg_batch_get_contact_events": (C.c_int64, [vp, iptr, iptr, iptr, iptr, C.c_int64]),
        "rp3dg_world_update": (C.c_int, [vp, f]),
        "rp3dg_batch_create": (vp, [C.POINTER(vp), i32]),
        "rp3dg_batch_create_ex": (vp, [C.POINTER(vp), i32, C.POINTER(BatchOptions)]),
        "rp3dg_batch_destroy": (None, [vp]),
        "rp3dg_batch_step": (C.c_int, [vp, f, i32]),
        "rp3dg_batch_step_graph": (C.c_int, [vp, f, i32]),
        "rp3dg_batch_synchronize": (C.c_int, [vp]),
        "rp3dg_batch_snapshot": (C.c_int, [vp, C.POINTER(vp)]),
        "rp3dg_batch_restore": (C.c_int, [vp, vp]),
        "rp3dg_snapshot_bytes": (C.c_int64, [vp]),
        "rp3dg_snapshot_destroy": (None, [vp]),
        "rp3dg_batch_get_counts": (C.c_int, [vp, C.POINTER(Counts)]),
        "rp3dg_batch_get_bodies": (C.c_int, [vp, fptr, fptr, fptr, fptr]),
        "rp3dg_batch_set_bodies": (C.c_int, [vp, fptr, fptr, fptr, fptr]),
        "rp3dg_batch_get_pairs": (C.c_int, [vp, iptr, iptr, i64]),
        "rp3dg_batch_get_contacts": (C.c_int, [vp, C.POINTER(ContactView), i64]),
        "rp3dg_batch_get_joint_order": (C.c_int, [vp, iptr, i64]),
        "rp3dg_batch_profile": (C.c_int, [vp, i32, C.c_char_p]),
        "rp3dg_batch_get_kernel_times": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(i64)]),
        "rp3dg_batch_set_residual_mode": (C.c_int, [vp, f, i32]),
        "rp3dg_batch_get_iteration_stats": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                      C.POINTER(C.c_double)]),
        "rp3dg_batch_get_world_iterations": (C.c_int, [vp, iptr, fptr, i64]),
        "rp3dg_batch_get_warm_start_predictions": (i64, [vp, fptr, i64]),
        "rp3dg_mlp_gradient_noalloc": (C.c_int, [vp, vp, vp, i64, vp, vp]),
        "rp3dg_hull_create_from_points": (i32, [vp, vp, i32, i32, i32]),
        "rp3dg_collider_set_is_trigger": (C.c_int, [vp, i32, i32]),
        "rp3dg_collider_set_is_simulation_collider": (C.c_int, [vp, i32, i32]),
        "rp3dg_collider_set_is_world_query_collider": (C.c_int, [vp, i32, i32]),
        "rp3dg_collider_get_kind": (C.c_int, [vp, i32]),
        "rp3dg_batch_set_solver_shape": (C.c_int, [vp, i32]),
        "rp3dg_batch_body_set_transform": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_batch_body_set_velocity": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_batch_body_apply_force": (C.c_int, [vp, i32, fptr, fptr, i32]),
        "rp3dg_batch_body_apply_torque": (C.c_int, [vp, i32, fptr, i32]),
        "rp3dg_batch_body_reset_force_torque": (C.c_int, [vp, i32, i32, i32]),
        "rp3dg_batch_body_set_sleeping": (C.c_int, [vp, i32, i32]),
        "rp3dg_batch_get_force_torque": (C.c_int, [vp, fptr, fptr]),
        "rp3dg_batch_test_overlap_bodies": (C.c_int, [vp, i32, i32]),
        "rp3dg_batch_test_overlap": (C.c_int64, [vp, i32]),
        "rp3dg_batch_test_collision": (C.c_int64, [vp, i32, i32]),
        "rp3dg_batch_get_query_results": (C.c_int64, [vp, C.POINTER(ContactView), C.c_int64]),
        "rp3dg_world_upload": (vp, [vp]),
        "rp3dg_batch_set_group_order": (C.c_int, [vp, i32]),
        "rp3dg_world_set_contacts_position_correction": (C.c_int, [vp, i32]),
        "rp3dg_batch_set_contacts_position_correction": (C.c_int, [vp, i32]),
        "rp3dg_batch_raycast_all": (C.c_int64, [vp, C.c_int64, iptr, fptr, fptr, fptr, C.c_uint16, iptr, iptr, fptr, fptr,
                                                fptr, C.c_int64]),
        "rp3dg_body_set_transform": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_body_apply_force": (C.c_int, [vp, i32, fptr, fptr, i32]),
        "rp3dg_body_apply_torque": (C.c_int, [vp, i32, fptr, i32]),
        "rp3dg_body_reset_force_torque": (C.c_int, [vp, i32, i32, i32]),
        "rp3dg_body_get_force_torque": (C.c_int, [vp, i32, fptr, fptr]),
        "rp3dg_body_set_sleeping": (C.c_int, [vp, i32, i32]),
        "rp3dg_quickhull": (C.c_int, [vp, i32, i32, i32, fptr, i32, iptr, i32, iptr, i32, iptr, iptr]),
        "rp3dg_mlp_create": (vp, [C.c_uint64]),
        "rp3dg_mlp_destroy": (None, [vp]),
        "rp3dg_mlp_set_params": (C.c_int, [vp, fptr, i64]),
        "rp3dg_mlp_get_params": (C.c_int, [vp, fptr, i64]),
        "rp3dg_mlp_set_normalization": (C.c_int, [vp, fptr, fptr, fptr]),
        "rp3dg_mlp_forward": (C.c_int, [vp, fptr, i64, fptr]),
        "rp3dg_mlp_gradient": (C.c_int, [vp, fptr, fptr, i64, fptr, fptr]),
        "rp3dg_mlp_train": (C.c_int, [vp, fptr, fptr, i64, i32, i32, f, C.c_uint64, fptr]),
        "rp3dg_batch_set_warm_start": (C.c_int, [vp, i32, vp]),
        "rp3dg_batch_set_logging": (C.c_int, [vp, i64]),
        "rp3dg_batch_get_log": (i64, [vp, fptr, fptr, i64]),
        "rp3dg_batch_autotune": (C.c_int, [vp, f, i32, iptr, C.POINTER(C.c_double)]),
        "rp3dg_batch_get_tune_table": (i32, [vp, iptr, C.POINTER(C.c_double), i32]),
        "rp3dg_world_get_body_state": (C.c_int, [vp, i32, fptr, fptr, fptr, fptr]),
        "rp3dg_last_error": (C.c_char_p, []),
        "rp3dg_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().rp3dg_last_error().decode()


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
    return rc


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _f4(v):
    return (C.c_float * 4)(*[float(x) for x in v])


def world_from_scene(sc: S.Scene):
    """Creates a host world through the C ABI with the canonical creation sequence."""
    L = load()
    st = WorldSettings()
    L.rp3dg_world_settings_default(C.byref(st))
    s = sc.settings
    for k in range(3):
        st.gravity[k] = float(s["gravity"][k])
    st.persistent_contact_distance_threshold = float(s["persistent"])
    st.default_friction_coefficient = float(s["default_friction"])
    st.default_bounciness = float(s["default_bounciness"])
    st.restitution_velocity_threshold = float(s["restitution_threshold"])
    st.cos_angle_similar_contact_manifold = float(s["cos_similar"])
    st.default_velocity_solver_iterations = int(s["vel_iters"])
    st.default_position_solver_iterations = int(s["pos_iters"])
    st.is_sleeping_enabled = int(s["sleeping"])
    w = L.rp3dg_world_create(C.byref(st))
    points = getattr(sc, "hull_points", None) or [None] * len(getattr(sc, "hulls", []))
    for (verts, faces), pts in zip(getattr(sc, "hulls", []), points):
        if pts is not None:   # PhysicsCommon::createConvexMesh(VertexArray)
            p = np.ascontiguousarray(pts, np.float32)
            _check(L.rp3dg_hull_create_from_points(w, p.ctypes.data, len(p), 12, 0), "hull_create_from_points")
            continue
        v = np.ascontiguousarray(verts, np.float32)
        counts = np.array([len(f) for f in faces], np.int32)
        idx = np.array([i for f in faces for i in f], np.int32)
        _check(L.rp3dg_hull_create(w, _fp(v), len(v), _ip(counts), _ip(idx), len(faces)), "hull_create")
    for v, t, n in getattr(sc, "tri_meshes", []):
        _check(L.rp3dg_triangle_mesh_create(w, _fp(v), len(v), _ip(t), len(t), None if n is None else _fp(n), None),
               "triangle_mesh_create")
    for cols, rows, h in getattr(sc, "heightfields", []):
        _check(L.rp3dg_heightfield_create(w, cols, rows, _fp(h)), "heightfield_create")
    ci = 0
    cols = sc.colliders
    for bi, (bt, pos, q, lin, ang, grav, ld, ad, llock, alock) in enumerate(sc.bodies):
        b = _check(L.rp3dg_body_create(w, _f3(pos), _f4(q)), "body_create")
        _check(L.rp3dg_body_set_type(w, b, bt), "body_set_type")
        while ci < len(cols) and cols[ci][0] == bi:
            (_, shape, params, lpos, lq, fr, bo, de, cat, mask, kind) = cols[ci]
            c = _check(L.rp3dg_collider_add(w, b, shape, _f4(params), _f3(lpos), _f4(lq)), "collider_add")
            _check(L.rp3dg_collider_set_material(w, c, fr, bo, de), "set_material")
            if cat != 0x0001 or mask != 0xFFFF:
                _check(L.rp3dg_collider_set_filter(w, c, cat, mask), "set_filter")
            if kind & S.COLLIDER_TRIGGER:
                _check(L.rp3dg_collider_set_is_trigger(w, c, 1), "set_is_trigger")
            elif not kind & S.COLLIDER_SIMULATION:
                _check(L.rp3dg_collider_set_is_simulation_collider(w, c, 0), "set_is_simulation_collider")
            if not kind & S.COLLIDER_WORLD_QUERY:
                _check(L.rp3dg_collider_set_is_world_query_collider(w, c, 0), "set_is_world_query_collider")
            ci += 1
        if bt == S.DYNAMIC:
            _check(L.rp3dg_body_update_mass_properties(w, b), "update_mass_properties")
        if bt != S.STATIC:
            _check(L.rp3dg_body_set_velocity(w, b, _f3(lin), _f3(ang)), "set_velocity")
        if not grav:
            _check(L.rp3dg_body_enable_gravity(w, b, 0), "enable_gravity")
        if ld != 0.0 or ad != 0.0:
            _check(L.rp3dg_body_set_damping(w, b, ld, ad), "set_damping")
        if tuple(llock) != (1.0, 1.0, 1.0) or tuple(alock) != (1.0, 1.0, 1.0):
            _check(L.rp3dg_body_set_lock_factors(w, b, _f3(llock), _f3(alock)), "set_lock_factors")
    for (jt, b1, b2, anchor, axis, collide, flags, limits) in sc.joints:
        _check(L.rp3dg_joint_create(w, jt, b1, b2, _f3(anchor), _f3(axis), collide, flags, _f4(limits)),
               "joint_create")
    return w


def quickhull(points):
    """QuickHull polygons of a point cloud ([n, 3]) as PhysicsCommon::createConvexMesh(VertexArray) computes
    them: (vertices [nv, 3] float32, face vertex counts [nf], concatenated face indices)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = len(pts)
    cap_v, cap_f, cap_i = max(4, n), max(4, 2 * n), max(12, 6 * n)
    v = np.zeros((cap_v, 3), np.float32)
    c = np.zeros(cap_f, np.int32)
    i = np.zeros(cap_i, np.int32)
    nv, nf = C.c_int32(), C.c_int32()
    rc = load().rp3dg_quickhull(pts.ctypes.data, n, 12, 0, _fp(v), cap_v, _ip(c), cap_f, _ip(i), cap_i, C.byref(nv),
                                C.byref(nf))
    _check(rc, "quickhull")
    return v[:nv.value].copy(), c[:nf.value].copy(), i[:int(c[:nf.value].sum())].copy()


class Snapshot:
    """A batch's state held in device memory (freed with the object)."""

    def __init__(self, h):
        self.h = h

    @property
    def nbytes(self) -> int:
        return int(load().rp3dg_snapshot_bytes(self.h))

    def close(self):
        if self.h:
            load().rp3dg_snapshot_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """One or many worlds stepped together on the GPU."""

    def __init__(self, scenes, solver="auto", fused_block=0, pairs_per_collider=0, group_bodies=0, broadphase="auto"):
        L = load()
        if isinstance(scenes, S.Scene):
            scenes = [scenes]
        self.scenes = list(scenes)
        self.worlds = [world_from_scene(sc) for sc in self.scenes]
        arr = (C.c_void_p * len(self.worlds))(*self.worlds)
        opt = BatchOptions()
        opt.solver = SOLVER_MODES[solver]
        opt.fused_block = int(fused_block)
        opt.pairs_per_collider = int(pairs_per_collider)
        opt.group_bodies = int(group_bodies)
        opt.broadphase = {"auto": 0, "tree": 1, "world_scan": 2}[broadphase]
        self.h = L.rp3dg_batch_create_ex(arr, len(self.worlds), C.byref(opt))
        if not self.h:
            raise RuntimeError("rp3dg_batch_create failed: " + last_error())
        self.nb = sum(sc.nb_bodies for sc in self.scenes)
        self.body_base = np.cumsum([0] + [sc.nb_bodies for sc in self.scenes])
        self.collider_base = np.cumsum([0] + [len(sc.colliders) for sc in self.scenes])
        self.joint_base = np.cumsum([0] + [len(sc.joints) for sc in self.scenes])

    def close(self):
        L = load()
        if self.h:
            L.rp3dg_batch_destroy(self.h)
            self.h = None
        for w in self.worlds:
            L.rp3dg_world_destroy(w)
        self.worlds = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt, n=1):
        _check(load().rp3dg_batch_step(self.h, float(dt), int(n)), "batch_step")

    def step_graph(self, dt, n=1):
        _check(load().rp3dg_batch_step_graph(self.h, float(dt), int(n)), "batch_step_graph")

    def sync(self):
        _check(load().rp3dg_batch_synchronize(self.h), "synchronize")

    def snapshot(self) -> "Snapshot":
        """The batch's state as a device copy (rp3dg_batch_snapshot)."""
        h = C.c_void_p()
        _check(load().rp3dg_batch_snapshot(self.h, C.byref(h)), "batch_snapshot")
        return Snapshot(h.value)

    def restore(self, snap: "Snapshot"):
        """Puts the batch back into the state of a snapshot taken from it (rp3dg_batch_restore)."""
        _check(load().rp3dg_batch_restore(self.h, snap.h), "batch_restore")

    def counts(self) -> dict:
        c = Counts()
        _check(load().rp3dg_batch_get_counts(self.h, C.byref(c)), "get_counts")
        out = {n: getattr(c, n) for n, _ in Counts._fields_ if n != "nb_tests"}
        out["nb_tests"] = list(c.nb_tests)
        return out

    def bodies(self):
        n = self.nb
        p = np.zeros((n, 3), np.float32)
        q = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        _check(load().rp3dg_batch_get_bodies(self.h, _fp(p), _fp(q), _fp(v), _fp(w)), "get_bodies")
        return p, q, v, w

    def set_contact_events(self, enable=True):
        """PhysicsWorld::setEventListener: record the contact events of every step from now on."""
        _check(load().rp3dg_batch_set_contact_events(self.h, int(enable)), "set_contact_events")

    def contact_events(self):
        """Events of the last step as an int32 array of rows (collider1, collider2, type, contact points),
        type 0 ContactStart, 1 ContactStay, 2 ContactExit; sorted by (collider1, collider2)."""
        L = load()
        cap = L.rp3dg_batch_get_contact_events(self.h, None, None, None, None, 0)
        if cap < 0:
            raise RuntimeError(f"get_contact_events: {last_error()}")
        out = np.zeros((4, max(1, cap)), np.int32)
        n = L.rp3dg_batch_get_contact_events(self.h, _ip(out[0]), _ip(out[1]), _ip(out[2]), _ip(out[3]), cap)
        if n < 0:
            raise RuntimeError(f"get_contact_events: {last_error()}")
        ev = out[:, :n].T.copy()
        return ev[np.lexsort((ev[:, 1], ev[:, 0]))] if n else ev

    def raycast(self, p1, p2, world=None, max_fraction=None, mask=0xFFFF):
        """PhysicsWorld::raycast with a closest-hit callback for n rays (p1, p2: [n, 3]; world: [n] world
        index, default 0).  Returns (collider [n], point [n, 3], normal [n, 3], fraction [n]) with batch
        collider indices, -1 for no hit."""
        p1 = np.ascontiguousarray(p1, np.float32).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, np.float32).reshape(-1, 3)
        n = len(p1)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        mf = None if max_fraction is None else np.ascontiguousarray(max_fraction, np.float32)
        col = np.zeros(n, np.int32)
        pt = np.zeros((n, 3), np.float32)
        nr = np.zeros((n, 3), np.float32)
        fr = np.zeros(n, np.float32)
        _check(load().rp3dg_batch_raycast(self.h, n, _ip(w) if w is not None else None, _fp(p1), _fp(p2),
                                          _fp(mf) if mf is not None else None, int(mask), _ip(col), _fp(pt), _fp(nr),
                                          _fp(fr)), "raycast")
        return col, pt, nr, fr

    def test_overlap_bodies(self, body1, body2):
        """PhysicsWorld::testOverlap(body1, body2) -> bool (batch body indices)."""
        r = load().rp3dg_batch_test_overlap_bodies(self.h, int(body1), int(body2))
        _check(r if r < 0 else 0, "test_overlap_bodies")
        return bool(r)

    def test_overlap(self, body=-1):
        """testOverlap(callback) (body < 0) or testOverlap(body, callback): overlapping (collider1,
        collider2, event type: 0 OverlapStart / 1 OverlapStay) arrays."""
        n = load().rp3dg_batch_test_overlap(self.h, int(body))
        _check(n if n < 0 else 0, "test_overlap")
        r = self.query_results(n)
        return r["colliders"][:, 0], r["colliders"][:, 1], r["event"]

    def test_collision(self, body1=-1, body2=-1):
        """testCollision(callback / body1 / body1, body2): one record per colliding pair (colliders,
        nb_points, normal_depth [4, 4], local1 / local2 [4, 3])."""
        n = load().rp3dg_batch_test_collision(self.h, int(body1), int(body2))
        _check(n if n < 0 else 0, "test_collision")
        return self.query_results(n)

    def query_results(self, n):
        cap = max(1, int(n))
        out = {
            "colliders": np.zeros((cap, 2), np.int32), "nb_points": np.zeros(cap, np.int32),
            "color": np.zeros(cap, np.int32), "normal_depth": np.zeros((cap, 4, 4), np.float32),
            "local1": np.zeros((cap, 4, 3), np.float32), "local2": np.zeros((cap, 4, 3), np.float32),
            "impulse": np.zeros((cap, 4), np.float32), "friction": np.zeros((cap, 3), np.float32),
        }
        v = ContactView(_ip(out["colliders"]), _ip(out["nb_points"]), _ip(out["color"]), _fp(out["normal_depth"]),
                        _fp(out["local1"]), _fp(out["local2"]), _fp(out["impulse"]), _fp(out["friction"]))
        k = load().rp3dg_batch_get_query_results(self.h, C.byref(v), cap)
        _check(k if k < 0 else 0, "get_query_results")
        res = {key: out[key][:k] for key in ("colliders", "nb_points", "normal_depth", "local1", "local2")}
        res["event"] = out["color"][:k]   # 0 Start, 1 Stay
        return res

    def raycast_all(self, p1, p2, world=None, max_fraction=None, mask=0xFFFF):
        """PhysicsWorld::raycast with a callback keeping every hit: (ray, collider, point, normal, fraction)
        arrays sorted by (ray, collider)."""
        L = load()
        p1 = np.ascontiguousarray(p1, np.float32).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, np.float32).reshape(-1, 3)
        n = len(p1)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        mf = None if max_fraction is None else np.ascontiguousarray(max_fraction, np.float32)
        args = (self.h, n, _ip(w) if w is not None else None, _fp(p1), _fp(p2), _fp(mf) if mf is not None else None,
                int(mask))
        total = L.rp3dg_batch_raycast_all(*args, None, None, None, None, None, 0)
        _check(total if total < 0 else 0, "raycast_all")
        cap = max(int(total), 1)
        ray = np.zeros(cap, np.int32)
        col = np.zeros(cap, np.int32)
        pt = np.zeros((cap, 3), np.float32)
        nr = np.zeros((cap, 3), np.float32)
        fr = np.zeros(cap, np.float32)
        k = L.rp3dg_batch_raycast_all(*args, _ip(ray), _ip(col), _fp(pt), _fp(nr), _fp(fr), cap)
        _check(k if k < 0 else 0, "raycast_all")
        return ray[:k], col[:k], pt[:k], nr[:k], fr[:k]

    def apply_forces(self, force=None, torque=None):
        """RigidBody::applyWorldForceAtCenterOfMass / applyWorldTorque on every body ([nb, 3] each, None =
        none) before the next step; the step clears them."""
        f = None if force is None else np.ascontiguousarray(force, np.float32)
        t = None if torque is None else np.ascontiguousarray(torque, np.float32)
        _check(load().rp3dg_batch_apply_forces(self.h, _fp(f) if f is not None else None,
                                               _fp(t) if t is not None else None), "apply_forces")

    def sleeping(self):
        """RigidBody::isSleeping of every body, int32 0 / 1."""
        out = np.zeros(self.nb, np.int32)
        _check(load().rp3dg_batch_get_sleeping(self.h, _ip(out), self.nb), "get_sleeping")
        return out

    def set_bodies(self, pos=None, quat=None, lin=None, ang=None):
        """RigidBody::setTransform / setLinearVelocity / setAngularVelocity for every body (None = keep)."""
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (pos, quat, lin, ang)]
        ptrs = [None if a is None else _fp(a) for a in arrs]
        _check(load().rp3dg_batch_set_bodies(self.h, *ptrs), "set_bodies")

    def body_ops(self, ops):
        """RigidBody calls between steps, in order: each op is (code, body, flags, a, c) with code one of
        OP_* (a = position / velocity / force / torque, c = orientation xyzw or application point, None
        where unused; flags FORCE_* for forces and torques, 1 = asleep for OP_SET_SLEEPING).  Queued and
        applied on the device before the next step or read-back."""
        L = load()

        def vec(x, n):
            return None if x is None else _fp(np.ascontiguousarray(x, np.float32).reshape(n))

        for code, body, flags, a, c in ops:
            body, flags = int(body), int(flags)
            if code == OP_SET_TRANSFORM:
                r = L.rp3dg_batch_body_set_transform(self.h, body, vec(a, 3), vec(c, 4))
            elif code == OP_SET_LINEAR_VELOCITY:
                r = L.rp3dg_batch_body_set_velocity(self.h, body, vec(a, 3), None)
            elif code == OP_SET_ANGULAR_VELOCITY:
                r = L.rp3dg_batch_body_set_velocity(self.h, body, None, vec(a, 3))
            elif code == OP_APPLY_FORCE:
                r = L.rp3dg_batch_body_apply_force(self.h, body, vec(a, 3), vec(c, 3), flags)
            elif code == OP_APPLY_TORQUE:
                r = L.rp3dg_batch_body_apply_torque(self.h, body, vec(a, 3), flags)
            elif code in (OP_RESET_FORCE, OP_RESET_TORQUE):
                r = L.rp3dg_batch_body_reset_force_torque(self.h, body, int(code == OP_RESET_FORCE),
                                                          int(code == OP_RESET_TORQUE))
            elif code == OP_SET_SLEEPING:
                r = L.rp3dg_batch_body_set_sleeping(self.h, body, flags)
            else:
                raise ValueError(f"unknown RigidBody call code {code}")
            _check(r, "body_ops")

    def force_torque(self):
        """RigidBody::getForce / getTorque of every body, ([nb, 3], [nb, 3])."""
        f = np.zeros((self.nb, 3), np.float32)
        t = np.zeros((self.nb, 3), np.float32)
        _check(load().rp3dg_batch_get_force_torque(self.h, _fp(f), _fp(t)), "get_force_torque")
        return f, t

    def pairs(self):
        cap = max(1, self.counts()["nb_pairs"])
        c1 = np.zeros(cap, np.int32)
        c2 = np.zeros(cap, np.int32)
        n = _check(load().rp3dg_batch_get_pairs(self.h, _ip(c1), _ip(c2), cap), "get_pairs")
        return c1[:n], c2[:n]

    def contacts(self):
        nm = max(1, self.counts()["nb_manifolds"])
        out = {
            "colliders": np.zeros((nm, 2), np.int32), "nb_points": np.zeros(nm, np.int32),
            "color": np.zeros(nm, np.int32), "normal_depth": np.zeros((nm, 4, 4), np.float32),
            "local1": np.zeros((nm, 4, 3), np.float32), "local2": np.zeros((nm, 4, 3), np.float32),
            "impulse": np.zeros((nm, 4), np.float32), "friction": np.zeros((nm, 3), np.float32),
            "manifold": np.zeros(nm, np.int32),
        }
        v = ContactView(_ip(out["colliders"]), _ip(out["nb_points"]), _ip(out["color"]), _fp(out["normal_depth"]),
                        _fp(out["local1"]), _fp(out["local2"]), _fp(out["impulse"]), _fp(out["friction"]),
                        _ip(out["manifold"]))
        n = _check(load().rp3dg_batch_get_contacts(self.h, C.byref(v), nm), "get_contacts")
        return {k: a[:n] for k, a in out.items()}

    def joint_order(self):
        n = max(1, self.counts()["nb_joints"])
        o = np.zeros(n, np.int32)
        k = _check(load().rp3dg_batch_get_joint_order(self.h, _ip(o), n), "get_joint_order")
        return o[:k]

    def profile(self, enable=True, only=None):
        _check(load().rp3dg_batch_profile(self.h, int(enable), only.encode() if only else None), "profile")

    def set_residual_mode(self, threshold, max_iterations=64):
        _check(load().rp3dg_batch_set_residual_mode(self.h, float(threshold), int(max_iterations)), "residual_mode")

    def iteration_stats(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(load().rp3dg_batch_get_iteration_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), "iteration_stats")
        return {"mean_iterations": a.value, "max_iterations": b.value, "max_residual": c.value}

    def world_iterations(self):
        """(velocity iterations [nworlds], last residual [nworlds]) of the last step."""
        nw = len(self.scenes)
        it = np.zeros(nw, np.int32)
        res = np.zeros(nw, np.float32)
        rc = load().rp3dg_batch_get_world_iterations(self.h, _ip(it), _fp(res), nw)
        _check(rc if rc < 0 else 0, "world_iterations")
        return it, res

    def warm_start_predictions(self):
        """Learned warm start predictions of the last step, [manifolds, 4, 3] in solver order."""
        L = load()
        n = L.rp3dg_batch_get_warm_start_predictions(self.h, None, 0)
        _check(n if n < 0 else 0, "warm_start_predictions")
        out = np.zeros((max(n, 1), 4, 3), np.float32)
        n = L.rp3dg_batch_get_warm_start_predictions(self.h, _fp(out), len(out))
        _check(n if n < 0 else 0, "warm_start_predictions")
        return out[:n]

    def autotune(self, dt, steps_per_trial=3):
        c = C.c_int32()
        ms = C.c_double()
        _check(load().rp3dg_batch_autotune(self.h, float(dt), int(steps_per_trial), C.byref(c), C.byref(ms)),
               "autotune")
        return c.value, ms.value

    def set_contacts_baumgarte(self, on=True):
        """PhysicsWorld::setContactsPositionCorrectionTechnique: BAUMGARTE_CONTACTS (on) or SPLIT_IMPULSES."""
        _check(load().rp3dg_batch_set_contacts_position_correction(self.h, 0 if on else 1), "contacts technique")

    def set_group_order(self, by_sweeps=True):
        _check(load().rp3dg_batch_set_group_order(self.h, int(by_sweeps)), "set_group_order")

    def set_solver_shape(self, choice):
        _check(load().rp3dg_batch_set_solver_shape(self.h, int(choice)), "set_solver_shape")

    def tune_table(self):
        """(choice, ms per step) of every trial of the last autotune call."""
        ch = (C.c_int32 * 16)()
        ms = (C.c_double * 16)()
        n = load().rp3dg_batch_get_tune_table(self.h, ch, ms, 16)
        _check(min(n, 0), "get_tune_table")
        return [(ch[i], ms[i]) for i in range(min(n, 16))]

    def set_warm_start(self, mode, mlp=None):
        m = {"plain": 0, "learned": 1}[mode]
        _check(load().rp3dg_batch_set_warm_start(self.h, m, mlp.h if mlp is not None else None), "set_warm_start")
        self._mlp = mlp   # keep the network alive while the batch may use it

    def set_logging(self, capacity):
        _check(load().rp3dg_batch_set_logging(self.h, int(capacity)), "set_logging")

    def get_log(self, capacity):
        f = np.zeros((capacity, N_FEATURES), np.float32)
        t = np.zeros((capacity, 3), np.float32)
        n = _check(load().rp3dg_batch_get_log(self.h, _fp(f), _fp(t), capacity), "get_log")
        return f[:n], t[:n]

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int64()
        _check(load().rp3dg_batch_get_kernel_times(self.h, name.encode(), C.byref(ms), C.byref(n)), "kernel_times")
        return ms.value, n.value


class Mlp:
    """The learned warm-start network (BF16 on the matrix cores); parameters live on the device."""

    def __init__(self, seed=1):
        self.h = load().rp3dg_mlp_create(int(seed))
        if not self.h:
            raise RuntimeError("rp3dg_mlp_create failed: " + last_error())

    def close(self):
        if self.h:
            load().rp3dg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        _check(load().rp3dg_mlp_get_params(self.h, _fp(p), MLP_PARAMS), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(load().rp3dg_mlp_set_params(self.h, _fp(p), p.size), "mlp_set_params")

    def set_normalization(self, mu, inv_sigma, out_scale):
        mu = np.ascontiguousarray(mu, np.float32)
        inv_sigma = np.ascontiguousarray(inv_sigma, np.float32)
        out_scale = np.ascontiguousarray(out_scale, np.float32)
        _check(load().rp3dg_mlp_set_normalization(self.h, _fp(mu), _fp(inv_sigma), _fp(out_scale)), "mlp_norm")

    def forward(self, feats):
        feats = np.ascontiguousarray(feats, np.float32)
        out = np.zeros((len(feats), 3), np.float32)
        _check(load().rp3dg_mlp_forward(self.h, _fp(feats), len(feats), _fp(out)), "mlp_forward")
        return out

    def gradient(self, feats, targets):
        feats = np.ascontiguousarray(feats, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(load().rp3dg_mlp_gradient(self.h, _fp(feats), _fp(targets), len(feats), _fp(g), C.byref(loss)),
               "mlp_gradient")
        return g, loss.value

    def train(self, feats, targets, epochs=10, minibatch=4096, lr=1e-3, seed=1):
        feats = np.ascontiguousarray(feats, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        loss = C.c_float()
        _check(load().rp3dg_mlp_train(self.h, _fp(feats), _fp(targets), len(feats), int(epochs), int(minibatch),
                                      float(lr), int(seed), C.byref(loss)), "mlp_train")
        return loss.value
