"""ORACLE / TEST INFRASTRUCTURE: ctypes binding of the reference harness.

Loads oracle/_ref/librp3d_ref_harness.so (the unmodified ReactPhysics3D
0.10.2 library compiled from its own sources by oracle/Makefile, plus
oracle/ref_harness.cpp).  Only tests, __graft_entry__.smoke() and the CPU
baseline leg of bench.py use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "librp3d_ref_harness.so")
REF_LIB_F64 = os.path.join(HERE, "_ref", "librp3d_ref_harness_f64.so")

_libs = {}


def available(f64=False) -> bool:
    return os.path.exists(REF_LIB_F64 if f64 else REF_LIB)


def load(f64=False):
    path = REF_LIB_F64 if f64 else REF_LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"reference harness not built: {path} (make -C oracle)")
    L = C.CDLL(path)
    vp, i32, f = C.c_void_p, C.c_int32, C.c_float
    fptr, iptr = C.POINTER(C.c_float), C.POINTER(C.c_int32)
    L.rp3dref_concave_overlapping_triangles.restype = i32
    L.rp3dref_concave_overlapping_triangles.argtypes = [vp, i32, fptr, fptr, iptr, fptr, fptr, i32]
    L.rp3dref_create.restype = vp
    L.rp3dref_create.argtypes = [vp, i32, vp, i32, vp, i32, vp, fptr, iptr]
    L.rp3dref_destroy.argtypes = [vp]
    L.rp3dref_step.argtypes = [vp, f]
    L.rp3dref_step_ordered.argtypes = [vp, f, iptr, iptr, i32]
    L.rp3dref_get_bodies.argtypes = [vp, fptr, fptr, fptr, fptr]
    L.rp3dref_get_mass_properties.argtypes = [vp, fptr, fptr, fptr]
    L.rp3dref_get_fat_aabbs.argtypes = [vp, fptr, fptr]
    L.rp3dref_get_broadphase_ids.argtypes = [vp, iptr]
    L.rp3dref_get_pairs.argtypes = [vp, iptr, iptr, i32]
    L.rp3dref_get_contacts.argtypes = [vp, iptr, iptr, fptr, fptr, fptr, fptr, fptr, i32]
    L.rp3dref_precision_bytes.restype = i32
    _libs[path] = L
    return L


def cpu_bench(sc, threads, warmup, steps, dt=1.0 / 60.0):
    """Times the reference PhysicsWorld::update() on `threads` copies of the scene, one per thread.
    Returns (seconds, contact points summed over the timed steps of all copies)."""
    L = load()
    L.rp3dref_bench.restype = C.c_int32
    L.rp3dref_bench.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                C.POINTER(C.c_double)]
    from reactphysics3d_amd.scene import JOINT_DTYPE
    s = np.array(sc.settings)
    b = sc.body_array()
    c = sc.collider_array()
    j = sc.joint_array() if sc.joints else np.zeros(1, dtype=JOINT_DTYPE)
    hv, hi = (sc.hull_arrays() if getattr(sc, "hulls", None) else (None, None))
    out = (C.c_double * 2)()
    rc = L.rp3dref_bench(s.ctypes.data, len(b), b.ctypes.data, len(c), c.ctypes.data, len(sc.joints), j.ctypes.data,
                         _fp(hv) if hv is not None else None, _ip(hi) if hi is not None else None,
                         threads, warmup, steps, float(dt), out)
    if rc != 0:
        raise RuntimeError("rp3dref_bench failed")
    return out[0], out[1]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class RefWorld:
    """A reference PhysicsWorld built from a reactphysics3d_amd.scene.Scene."""

    def __init__(self, sc, f64=False):
        self.L = load(f64)
        self.sc = sc
        self._s = np.array(sc.settings)
        self._b = sc.body_array()
        self._c = sc.collider_array()
        self._j = sc.joint_array() if sc.joints else np.zeros(1, dtype=self._jdtype())
        hv, hi = sc.hull_arrays()
        self._hv, self._hi = hv, hi
        self.h = self.L.rp3dref_create(self._s.ctypes.data, len(self._b), self._b.ctypes.data, len(self._c),
                                       self._c.ctypes.data, len(sc.joints), self._j.ctypes.data,
                                       _fp(hv) if hv is not None else None, _ip(hi) if hi is not None else None)
        if not self.h:
            raise RuntimeError("rp3dref_create failed (unsupported shape?)")
        self.nb = len(self._b)
        self.nc = len(self._c)

    @staticmethod
    def _jdtype():
        from reactphysics3d_amd.scene import JOINT_DTYPE
        return JOINT_DTYPE

    def close(self):
        if self.h:
            self.L.rp3dref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt):
        self.L.rp3dref_step(self.h, float(dt))

    def set_joint_order(self, order):
        o = np.ascontiguousarray(order, np.int32)
        self.L.rp3dref_set_joint_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]
        if self.L.rp3dref_set_joint_order(self.h, _ip(o), len(o)) != 0:
            raise RuntimeError("set_joint_order: unsupported joint kind")

    def enable_events(self):
        """PhysicsWorld::setEventListener with a recorder of the contact events of every step."""
        self.L.rp3dref_enable_events.argtypes = [C.c_void_p]
        self.L.rp3dref_enable_events(self.h)

    def events(self):
        """Contact events of the last step, rows (collider1, collider2, type, contact points) sorted by
        (collider1, collider2); type 0 ContactStart, 1 ContactStay, 2 ContactExit."""
        cap = 64 * max(1, self.nc)
        out = np.zeros((cap, 4), np.int32)
        self.L.rp3dref_get_events.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]
        n = self.L.rp3dref_get_events(self.h, _ip(out), cap)
        assert n >= 0
        ev = out[:n]
        return ev[np.lexsort((ev[:, 1], ev[:, 0]))] if n else ev

    def raycast(self, p1, p2, max_fraction=None, mask=0xFFFF):
        """PhysicsWorld::raycast with a closest-hit callback for each ray (p1, p2: [n, 3]).  Returns
        (collider [n], point [n, 3], normal [n, 3], fraction [n]); collider -1 for no hit."""
        p1 = np.ascontiguousarray(p1, np.float32).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, np.float32).reshape(-1, 3)
        n = len(p1)
        mf = None if max_fraction is None else np.ascontiguousarray(max_fraction, np.float32)
        col = np.zeros(n, np.int32)
        pt = np.zeros((n, 3), np.float32)
        nr = np.zeros((n, 3), np.float32)
        fr = np.zeros(n, np.float32)
        fp = C.POINTER(C.c_float)
        self.L.rp3dref_raycast.argtypes = [C.c_void_p, C.c_int32, fp, fp, fp, C.c_int32, C.POINTER(C.c_int32), fp, fp, fp]
        self.L.rp3dref_raycast(self.h, n, _fp(p1), _fp(p2), _fp(mf) if mf is not None else None, int(mask), _ip(col),
                               _fp(pt), _fp(nr), _fp(fr))
        return col, pt, nr, fr


    def set_contacts_baumgarte(self, on=True):
        """PhysicsWorld::setContactsPositionCorrectionTechnique(BAUMGARTE_CONTACTS or SPLIT_IMPULSES)."""
        self.L.rp3dref_set_contacts_baumgarte.argtypes = [C.c_void_p, C.c_int]
        self.L.rp3dref_set_contacts_baumgarte(self.h, int(bool(on)))

    def test_overlap_bodies(self, b1, b2):
        """PhysicsWorld::testOverlap(body1, body2) -> bool."""
        self.L.rp3dref_test_overlap_bodies.argtypes = [C.c_void_p, C.c_int, C.c_int]
        self.L.rp3dref_test_overlap_bodies.restype = C.c_int
        return bool(self.L.rp3dref_test_overlap_bodies(self.h, int(b1), int(b2)))

    def test_overlap(self, body=-1):
        """testOverlap(callback) (body < 0) or testOverlap(body, callback): (collider1, collider2, event
        type) arrays in the reference's report order."""
        ip = C.POINTER(C.c_int32)
        self.L.rp3dref_test_overlap.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, C.c_int]
        self.L.rp3dref_test_overlap.restype = C.c_int
        cap = 64 * max(1, self.nc)
        c1 = np.zeros(cap, np.int32)
        c2 = np.zeros(cap, np.int32)
        ev = np.zeros(cap, np.int32)
        n = self.L.rp3dref_test_overlap(self.h, int(body), _ip(c1), _ip(c2), _ip(ev), cap)
        if n > cap:   # a second call would not see the same state: the query changes it
            raise RuntimeError(f"test_overlap: {n} pairs, capacity {cap}")
        return c1[:n], c2[:n], ev[:n]

    def test_collision(self, b1=-1, b2=-1):
        """testCollision(callback / body1 / body1, body2): contact records as Batch.test_collision."""
        ip = C.POINTER(C.c_int32)
        fp = C.POINTER(C.c_float)
        self.L.rp3dref_test_collision.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, ip, fp, fp, fp, C.c_int]
        self.L.rp3dref_test_collision.restype = C.c_int
        cap = 64 * max(1, self.nc)
        out = {"colliders": np.zeros((cap, 2), np.int32), "nb_points": np.zeros(cap, np.int32),
               "event": np.zeros(cap, np.int32),
               "normal_depth": np.zeros((cap, 4, 4), np.float32), "local1": np.zeros((cap, 4, 3), np.float32),
               "local2": np.zeros((cap, 4, 3), np.float32)}
        n = self.L.rp3dref_test_collision(self.h, int(b1), int(b2), _ip(out["colliders"]), _ip(out["nb_points"]),
                                          _ip(out["event"]), _fp(out["normal_depth"]), _fp(out["local1"]),
                                          _fp(out["local2"]), cap)
        if n > cap:
            raise RuntimeError(f"test_collision: {n} contact pairs, capacity {cap}")
        return {k: a[:n] for k, a in out.items()}

    def broadphase_only(self):
        """computeBroadPhase alone (the broadphase every query runs in its world)."""
        self.L.rp3dref_broadphase_only.argtypes = [C.c_void_p]
        self.L.rp3dref_broadphase_only(self.h)

    def raycast_all(self, p1, p2, max_fraction=None, mask=0xFFFF):
        """Every hit (callback returning 1), in the reference's report order: (ray, collider, point,
        normal, fraction) arrays."""
        p1 = np.ascontiguousarray(p1, np.float32)
        p2 = np.ascontiguousarray(p2, np.float32)
        n = len(p1)
        mf = None if max_fraction is None else np.ascontiguousarray(max_fraction, np.float32)
        fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
        self.L.rp3dref_raycast_all.argtypes = [C.c_void_p, C.c_int, fp, fp, fp, C.c_int, ip, fp, C.c_int]
        self.L.rp3dref_raycast_all.restype = C.c_int
        cap = 0
        while True:
            rc = np.zeros((max(cap, 1), 2), np.int32)
            pnf = np.zeros((max(cap, 1), 7), np.float32)
            total = self.L.rp3dref_raycast_all(self.h, n, _fp(p1), _fp(p2), _fp(mf) if mf is not None else None, int(mask),
                                               rc.ctypes.data_as(ip), _fp(pnf), cap)
            if total <= cap:
                return rc[:total, 0], rc[:total, 1], pnf[:total, 0:3], pnf[:total, 3:6], pnf[:total, 6]
            cap = total
    def apply_forces(self, force=None, torque=None):
        """RigidBody::applyWorldForceAtCenterOfMass / applyWorldTorque on every body ([nb, 3] each)."""
        f = None if force is None else np.ascontiguousarray(force, np.float32)
        t = None if torque is None else np.ascontiguousarray(torque, np.float32)
        fp = C.POINTER(C.c_float)
        self.L.rp3dref_apply_forces.argtypes = [C.c_void_p, fp, fp]
        self.L.rp3dref_apply_forces(self.h, _fp(f) if f is not None else None, _fp(t) if t is not None else None)


    def body_ops(self, ops):
        """RigidBody calls on the reference world, in order: (code, body, flags, a, c) as the device
        library's Batch.body_ops takes them."""
        fp = C.POINTER(C.c_float)
        self.L.rp3dref_body_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, fp, fp]
        self.L.rp3dref_body_op.restype = C.c_int
        for code, body, flags, a, c in ops:
            av = None if a is None else np.ascontiguousarray(a, np.float32)
            cv = None if c is None else np.ascontiguousarray(c, np.float32)
            r = self.L.rp3dref_body_op(self.h, int(body), int(code), int(flags),
                                       _fp(av) if av is not None else None, _fp(cv) if cv is not None else None)
            if r != 0:
                raise RuntimeError(f"reference RigidBody call {code} on body {body} failed")

    def force_torque(self):
        n = self.nb
        f = np.zeros((n, 3), np.float32)
        t = np.zeros((n, 3), np.float32)
        fp = C.POINTER(C.c_float)
        self.L.rp3dref_get_force_torque.argtypes = [C.c_void_p, fp, fp]
        self.L.rp3dref_get_force_torque(self.h, _fp(f), _fp(t))
        return f, t
    def sleeping(self):
        """RigidBody::isSleeping of every body, as int32 0 / 1."""
        out = np.zeros(self.nb, np.int32)
        self.L.rp3dref_get_sleeping.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        self.L.rp3dref_get_sleeping(self.h, _ip(out))
        return out

    def step_ordered(self, dt, order_c1, order_c2, order_sub=None):
        """order_sub: index of each manifold among its pair's (a convex vs concave pair has up to four),
        carried in bits 28-31 of collider 2."""
        o1 = np.ascontiguousarray(order_c1, np.int32)
        o2 = np.ascontiguousarray(order_c2, np.int32)
        if order_sub is not None:
            o2 = np.ascontiguousarray(o2 | (np.asarray(order_sub, np.int32) << 28), np.int32)
        rc = self.L.rp3dref_step_ordered(self.h, float(dt), _ip(o1), _ip(o2), len(o1))
        if rc != 0:
            raise RuntimeError("coloured-order step: a reference manifold is missing from the given order")

    def step_ordered_mode(self, dt, order_c1, order_c2, residual=0.0, max_iterations=0, pred=None):
        """Ordered step with the residual stopping rule (residual > 0, at most max_iterations sweeps)
        and / or the learned warm start (pred: [manifolds, 12] in the given order).  Returns
        (sweeps run, last residual)."""
        o1 = np.ascontiguousarray(order_c1, np.int32)
        o2 = np.ascontiguousarray(order_c2, np.int32)
        pr = None if pred is None else np.ascontiguousarray(np.asarray(pred, np.float32).reshape(-1, 12))
        it = C.c_int32()
        res = C.c_float()
        fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
        self.L.rp3dref_step_ordered_mode.argtypes = [C.c_void_p, C.c_float, ip, ip, C.c_int32, C.c_float, C.c_int32,
                                                     fp, ip, fp]
        rc = self.L.rp3dref_step_ordered_mode(self.h, float(dt), _ip(o1), _ip(o2), len(o1), float(residual),
                                              int(max_iterations), _fp(pr) if pr is not None else None,
                                              C.byref(it), C.byref(res))
        if rc != 0:
            raise RuntimeError("coloured-order step: a reference manifold is missing from the given order")
        return it.value, res.value

    def set_iterations(self, velocity, position):
        """PhysicsWorld::setNbIterationsVelocitySolver / setNbIterationsPositionSolver."""
        self.L.rp3dref_set_iterations.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        self.L.rp3dref_set_iterations(self.h, int(velocity), int(position))

    def bodies(self):
        p = np.zeros((self.nb, 3), np.float32)
        q = np.zeros((self.nb, 4), np.float32)
        v = np.zeros((self.nb, 3), np.float32)
        w = np.zeros((self.nb, 3), np.float32)
        self.L.rp3dref_get_bodies(self.h, _fp(p), _fp(q), _fp(v), _fp(w))
        return p, q, v, w

    def mass_properties(self):
        m = np.zeros(self.nb, np.float32)
        i = np.zeros((self.nb, 3), np.float32)
        c = np.zeros((self.nb, 3), np.float32)
        self.L.rp3dref_get_mass_properties(self.h, _fp(m), _fp(i), _fp(c))
        return m, i, c

    def fat_aabbs(self):
        mn = np.zeros((self.nc, 3), np.float32)
        mx = np.zeros((self.nc, 3), np.float32)
        self.L.rp3dref_get_fat_aabbs(self.h, _fp(mn), _fp(mx))
        return mn, mx

    def broadphase_ids(self):
        ids = np.zeros(self.nc, np.int32)
        self.L.rp3dref_get_broadphase_ids(self.h, _ip(ids))
        return ids

    def concave_overlapping_triangles(self, collider, aabb_min, aabb_max, capacity=4096):
        """ConcaveShape::computeOverlappingTriangles of a concave collider for an AABB in its local space:
        (shape ids, vertices [n, 3, 3], vertex normals [n, 3, 3]) in the middle phase's order."""
        ids = np.zeros(capacity, np.int32)
        vs = np.zeros((capacity, 3, 3), np.float32)
        ns = np.zeros((capacity, 3, 3), np.float32)
        lo = np.ascontiguousarray(aabb_min, np.float32)
        hi = np.ascontiguousarray(aabb_max, np.float32)
        n = self.L.rp3dref_concave_overlapping_triangles(self.h, collider, _fp(lo), _fp(hi), _ip(ids), _fp(vs),
                                                         _fp(ns), capacity)
        if n < 0 or n > capacity:
            raise RuntimeError("rp3dref_concave_overlapping_triangles failed")
        return ids[:n], vs[:n], ns[:n]

    def pairs(self):
        cap = 64 * max(1, self.nc)
        c1 = np.zeros(cap, np.int32)
        c2 = np.zeros(cap, np.int32)
        n = self.L.rp3dref_get_pairs(self.h, _ip(c1), _ip(c2), cap)
        assert n >= 0
        return c1[:n], c2[:n]

    def contacts(self):
        cap = 64 * max(1, self.nc)
        out = {
            "colliders": np.zeros((cap, 2), np.int32), "nb_points": np.zeros(cap, np.int32),
            "normal_depth": np.zeros((cap, 4, 4), np.float32), "local1": np.zeros((cap, 4, 3), np.float32),
            "local2": np.zeros((cap, 4, 3), np.float32), "impulse": np.zeros((cap, 4), np.float32),
            "friction": np.zeros((cap, 3), np.float32),
        }
        n = self.L.rp3dref_get_contacts(self.h, _ip(out["colliders"]), _ip(out["nb_points"]),
                                        _fp(out["normal_depth"]), _fp(out["local1"]), _fp(out["local2"]),
                                        _fp(out["impulse"]), _fp(out["friction"]), cap)
        assert n >= 0
        out = {k: a[:n] for k, a in out.items()}
        # index of each manifold among its pair's (created one after the other)
        seen = {}
        sub = np.zeros(n, np.int32)
        for m, (a, b) in enumerate(out["colliders"]):
            sub[m] = seen.get((int(a), int(b)), 0)
            seen[(int(a), int(b))] = sub[m] + 1
        out["manifold"] = sub
        return out


class RefBatch:
    """Reference PhysicsWorlds, one per scene, stepped together in the coloured order on host threads
    (rp3dref_batch_step_ordered) and read back concatenated in batch numbering."""

    def __init__(self, scenes, threads=None):
        self.worlds = [RefWorld(sc) for sc in scenes]
        self.L = self.worlds[0].L
        self.nw = len(self.worlds)
        self.h = (C.c_void_p * self.nw)(*[w.h for w in self.worlds])
        nb = np.array([w.nb for w in self.worlds], np.int64)
        nc = np.array([w.nc for w in self.worlds], np.int64)
        self.body_base = np.concatenate([[0], np.cumsum(nb)])
        self.collider_base = np.concatenate([[0], np.cumsum(nc)]).astype(np.int32)
        self.threads = int(threads or min(16, os.cpu_count() or 1))
        L = self.L
        vpp = C.POINTER(C.c_void_p)
        fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
        lp = C.POINTER(C.c_int64)
        L.rp3dref_batch_step_ordered.restype = C.c_int32
        L.rp3dref_batch_step_ordered.argtypes = [vpp, C.c_int32, C.c_float, ip, ip, lp, C.c_float, C.c_int32, fp,
                                                 ip, fp, C.c_int32]
        L.rp3dref_batch_get_bodies.argtypes = [vpp, C.c_int32, fp, fp, fp, fp]
        L.rp3dref_batch_get_pairs.restype = C.c_int64
        L.rp3dref_batch_get_pairs.argtypes = [vpp, C.c_int32, ip, ip, ip, C.c_int64]
        L.rp3dref_batch_get_contacts.restype = C.c_int64
        L.rp3dref_batch_get_contacts.argtypes = [vpp, C.c_int32, ip, ip, ip, fp, fp, fp, fp, fp, C.c_int64]

    def close(self):
        for w in self.worlds:
            w.close()

    def set_joint_orders(self, orders):
        for w, o in zip(self.worlds, orders):
            w.set_joint_order(o)

    def step_ordered(self, dt, c1, c2, world_of_manifold, residual=0.0, max_iterations=0, pred=None):
        """One ordered step of every world.  c1 / c2: batch-global collider indices of the device's
        manifolds in solve order (each world's manifolds contiguous or not: the order within a world is
        kept); world_of_manifold: world of each.  pred: [n, 12] learned warm start predictions of those
        manifolds, or None.  Returns (iterations [nw], last residual [nw])."""
        wm = np.asarray(world_of_manifold, np.int64)
        order = np.argsort(wm, kind="stable")
        cb = self.collider_base.astype(np.int64)
        o1 = np.ascontiguousarray(np.asarray(c1, np.int64)[order] - cb[wm[order]], np.int32)
        o2 = np.ascontiguousarray(np.asarray(c2, np.int64)[order] - cb[wm[order]], np.int32)
        moff = np.ascontiguousarray(np.concatenate([[0], np.cumsum(np.bincount(wm, minlength=self.nw))]), np.int64)
        pr = None
        if pred is not None:
            pr = np.ascontiguousarray(np.asarray(pred, np.float32).reshape(-1, 12)[order])
        iters = np.zeros(self.nw, np.int32)
        res = np.zeros(self.nw, np.float32)
        rc = self.L.rp3dref_batch_step_ordered(self.h, self.nw, float(dt), _ip(o1), _ip(o2),
                                               moff.ctypes.data_as(C.POINTER(C.c_int64)), float(residual),
                                               int(max_iterations), _fp(pr) if pr is not None else None, _ip(iters),
                                               _fp(res), self.threads)
        if rc != 0:
            raise RuntimeError(f"coloured-order step: world {-rc - 1} has a manifold missing from the given order")
        return iters, res

    def bodies(self):
        n = int(self.body_base[-1])
        p = np.zeros((n, 3), np.float32)
        q = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        self.L.rp3dref_batch_get_bodies(self.h, self.nw, _fp(p), _fp(q), _fp(v), _fp(w))
        return p, q, v, w

    def pairs(self):
        cap = 64 * int(self.collider_base[-1])
        c1 = np.zeros(cap, np.int32)
        c2 = np.zeros(cap, np.int32)
        n = self.L.rp3dref_batch_get_pairs(self.h, self.nw, _ip(self.collider_base), _ip(c1), _ip(c2), cap)
        assert n >= 0
        return c1[:n], c2[:n]

    def contacts(self):
        cap = 16 * int(self.collider_base[-1])
        out = {
            "colliders": np.zeros((cap, 2), np.int32), "nb_points": np.zeros(cap, np.int32),
            "normal_depth": np.zeros((cap, 4, 4), np.float32), "local1": np.zeros((cap, 4, 3), np.float32),
            "local2": np.zeros((cap, 4, 3), np.float32), "impulse": np.zeros((cap, 4), np.float32),
            "friction": np.zeros((cap, 3), np.float32),
        }
        n = self.L.rp3dref_batch_get_contacts(self.h, self.nw, _ip(self.collider_base), _ip(out["colliders"]),
                                              _ip(out["nb_points"]), _fp(out["normal_depth"]), _fp(out["local1"]),
                                              _fp(out["local2"]), _fp(out["impulse"]), _fp(out["friction"]), cap)
        assert n >= 0
        return {k: a[:n] for k, a in out.items()}


def quickhull(points):
    """The reference's QuickHull::computeConvexHull of a point cloud ([n, 3] float32): (vertices [nv, 3],
    face vertex counts [nf], concatenated face indices), or None when the reference refuses the points."""
    L = load()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = len(pts)
    cap_v, cap_f, cap_i = max(4, n), max(4, 2 * n), max(12, 6 * n)
    v = np.zeros((cap_v, 3), np.float32)
    c = np.zeros(cap_f, np.int32)
    i = np.zeros(cap_i, np.int32)
    nv, nf = C.c_int32(), C.c_int32()
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
    L.rp3dref_quickhull.argtypes = [fp, C.c_int32, fp, C.c_int32, ip, C.c_int32, ip, C.c_int32, ip, ip]
    rc = L.rp3dref_quickhull(_fp(pts), n, _fp(v), cap_v, _ip(c), cap_f, _ip(i), cap_i, C.byref(nv), C.byref(nf))
    if rc == -1:
        return None
    assert rc == 0
    return v[:nv.value].copy(), c[:nf.value].copy(), i[:int(c[:nf.value].sum())].copy()
