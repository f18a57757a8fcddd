"""Seeded scene descriptions for the step path.

A Scene is a flat description (structured numpy arrays) of one PhysicsWorld:
settings, bodies, colliders (grouped by body, in creation order) and joints.
The same description is loaded into the MI355X library through its C ABI
(``reactphysics3d_amd.lib.DeviceWorld``) and, in tests, into the reference
harness, always with the same creation sequence (body, setType, colliders,
materials, mass properties, velocities), so both start from identical state.

Random numbers come from a project-owned splitmix64 stream (no
``std::*_distribution`` / numpy Generator behaviour is involved): floats are
the top 24 bits of each 64-bit output, quaternions use Shoemake's method and
normals Box-Muller.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
SPHERE, CAPSULE, BOX, HULL = 0, 1, 2, 3
# concave shapes (static or kinematic bodies only); params = (data index, scale x, y, z)
CONCAVE_MESH, HEIGHTFIELD = 4, 5
BALLSOCKET, SLIDER, HINGE, FIXED = 0, 1, 2, 3
# Joint flag bits (rp3dg_joint_create): limit, motor, Baumgarte position correction (RP3DG_JOINT_BAUMGARTE)
JOINT_LIMIT, JOINT_MOTOR, JOINT_BAUMGARTE = 1, 2, 4

SETTINGS_DTYPE = np.dtype([
    ("gravity", "<f4", 3), ("vel_iters", "<i4"), ("pos_iters", "<i4"),
    ("persistent", "<f4"), ("restitution_threshold", "<f4"), ("cos_similar", "<f4"),
    ("default_friction", "<f4"), ("default_bounciness", "<f4"), ("sleeping", "<i4"),
])
BODY_DTYPE = np.dtype([
    ("type", "<i4"), ("pos", "<f4", 3), ("quat", "<f4", 4), ("lin", "<f4", 3), ("ang", "<f4", 3),
    ("gravity", "<i4"), ("ldamp", "<f4"), ("adamp", "<f4"), ("llock", "<f4", 3), ("alock", "<f4", 3),
])
# Collider kind bits (RP3DG_COLLIDER_*): Collider::getIsTrigger / getIsSimulationCollider / getIsWorldQueryCollider
COLLIDER_TRIGGER, COLLIDER_SIMULATION, COLLIDER_WORLD_QUERY = 1, 2, 4
COLLIDER_DTYPE = np.dtype([
    ("body", "<i4"), ("shape", "<i4"), ("params", "<f4", 4), ("lpos", "<f4", 3), ("lquat", "<f4", 4),
    ("friction", "<f4"), ("bounciness", "<f4"), ("density", "<f4"), ("category", "<i4"), ("mask", "<i4"),
    ("kind", "<i4"),
])
JOINT_DTYPE = np.dtype([
    ("type", "<i4"), ("body1", "<i4"), ("body2", "<i4"), ("anchor", "<f4", 3), ("axis", "<f4", 3),
    ("collide", "<i4"), ("flags", "<i4"), ("limits", "<f4", 4),
])


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014); floats from the top 24 bits."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform01(self) -> float:
        return (self.next_u64() >> 40) * (1.0 / 16777216.0)

    def uniform(self, a: float, b: float) -> float:
        return a + (b - a) * self.uniform01()

    def normal(self) -> float:
        u1 = max(self.uniform01(), 1.0 / 16777216.0)
        u2 = self.uniform01()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def quaternion(self):
        """Uniform random rotation (Shoemake 1992), returned as (x, y, z, w)."""
        u1, u2, u3 = self.uniform01(), self.uniform01(), self.uniform01()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
             b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
        n = math.sqrt(sum(c * c for c in q))
        return tuple(c / n for c in q)


@dataclass
class Scene:
    name: str
    settings: np.ndarray
    bodies: list = field(default_factory=list)
    colliders: list = field(default_factory=list)
    joints: list = field(default_factory=list)
    hulls: list = field(default_factory=list)   # (vertices [V, 3] float32, faces: list of CCW index lists)
    hull_points: list = field(default_factory=list)   # per hull: the point cloud it was built from, or None
    tri_meshes: list = field(default_factory=list)    # (vertices [V, 3] float32, triangles [T, 3] int32, normals or None)
    heightfields: list = field(default_factory=list)  # (columns, rows, heights [rows, columns] float32)

    # -- builders -------------------------------------------------------------------------------
    def add_body(self, btype, pos, quat=(0.0, 0.0, 0.0, 1.0), lin=(0.0, 0.0, 0.0), ang=(0.0, 0.0, 0.0),
                 gravity=True, ldamp=0.0, adamp=0.0, llock=(1.0, 1.0, 1.0), alock=(1.0, 1.0, 1.0)):
        """llock / alock: RigidBody::setLinearLockAxisFactor / setAngularLockAxisFactor."""
        self.bodies.append((btype, tuple(pos), tuple(quat), tuple(lin), tuple(ang), int(gravity), ldamp, adamp,
                            tuple(llock), tuple(alock)))
        return len(self.bodies) - 1

    def add_collider(self, body, shape, params, lpos=(0.0, 0.0, 0.0), lquat=(0.0, 0.0, 0.0, 1.0),
                     friction=None, bounciness=None, density=1.0, category=0x0001, mask=0xFFFF, trigger=False,
                     simulation=None, world_query=True):
        """trigger / simulation / world_query: Collider::setIsTrigger / setIsSimulationCollider /
        setIsWorldQueryCollider (a trigger is not a simulation collider unless simulation=True is asked,
        which the reference turns back into a plain simulation collider)."""
        if self.colliders and self.colliders[-1][0] > body:
            raise ValueError("colliders must be added in body order")
        p = list(params) + [0.0] * (4 - len(params))
        f = float(self.settings["default_friction"]) if friction is None else friction
        b = float(self.settings["default_bounciness"]) if bounciness is None else bounciness
        sim = (not trigger) if simulation is None else bool(simulation)
        kind = (COLLIDER_TRIGGER if trigger and not sim else 0) | (COLLIDER_SIMULATION if sim else 0) | \
            (COLLIDER_WORLD_QUERY if world_query else 0)
        self.colliders.append((body, shape, tuple(p), tuple(lpos), tuple(lquat), f, b, density, category, mask, kind))
        return len(self.colliders) - 1

    def add_joint(self, jtype, body1, body2, anchor, axis=(0.0, 0.0, 1.0), collide=False, flags=0,
                  limits=(0.0, 0.0, 0.0, 0.0)):
        self.joints.append((jtype, body1, body2, tuple(anchor), tuple(axis), int(collide), flags, tuple(limits)))
        return len(self.joints) - 1

    def add_hull(self, vertices, faces):
        """A convex mesh (PhysicsCommon::createConvexMesh(PolygonVertexArray)); returns its index, the
        params[0] of a HULL collider."""
        self.hulls.append((np.ascontiguousarray(vertices, np.float32).reshape(-1, 3), [list(f) for f in faces]))
        self.hull_points.append(None)
        return len(self.hulls) - 1

    def add_hull_from_points(self, points):
        """A convex mesh from a point cloud (PhysicsCommon::createConvexMesh(VertexArray): QuickHull); the
        polygons are the ones the library's QuickHull computes, the reference's own."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        v, f = convex_hull(pts)
        self.hulls.append((v, f))
        self.hull_points.append(pts)
        return len(self.hulls) - 1

    def add_triangle_mesh(self, vertices, triangles, normals=None):
        """A triangle mesh (PhysicsCommon::createTriangleMesh(TriangleVertexArray)); returns its index, the
        params[0] of a CONCAVE_MESH collider."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self.tri_meshes.append((v, t, n))
        return len(self.tri_meshes) - 1

    def add_heightfield(self, heights):
        """A height field of float heights [rows, columns] (PhysicsCommon::createHeightField); returns its
        index, the params[0] of a HEIGHTFIELD collider."""
        h = np.ascontiguousarray(heights, np.float32)
        self.heightfields.append((h.shape[1], h.shape[0], h))
        return len(self.heightfields) - 1

    def hull_arrays(self):
        """(vertices float32 [sum V * 3], ints int32 [nb hulls, then per hull nv, nf, face sizes,
        face indices]) -- the layout the reference harness reads."""
        verts = []
        ints = [len(self.hulls)]
        for (v, faces), pts in zip(self.hulls, self.hull_points):
            if pts is not None:   # point cloud: the reference runs its own QuickHull (nf = -1)
                verts.append(pts.ravel())
                ints += [len(pts), -1]
                continue
            verts.append(v.ravel())
            ints += [len(v), len(faces)] + [len(f) for f in faces] + [i for f in faces for i in f]
        # concave shapes: triangle meshes, then height fields (heights padded to whole xyz triplets)
        ints.append(len(self.tri_meshes))
        for v, t, n in self.tri_meshes:
            ints += [len(v), len(t), 0 if n is None else 1] + [int(i) for i in t.ravel()]
            verts.append(v.ravel())
            if n is not None:
                verts.append(n.ravel())
        ints.append(len(self.heightfields))
        for cols, rows, h in self.heightfields:
            ints += [cols, rows]
            pad = (-h.size) % 3
            verts.append(np.concatenate([h.ravel(), np.zeros(pad, np.float32)]))
        vs = np.concatenate(verts) if verts else np.zeros(3, np.float32)
        return np.ascontiguousarray(vs, np.float32), np.array(ints, np.int32)

    # -- packed views ---------------------------------------------------------------------------
    def body_array(self) -> np.ndarray:
        return np.array(self.bodies, dtype=BODY_DTYPE)

    def collider_array(self) -> np.ndarray:
        return np.array(self.colliders, dtype=COLLIDER_DTYPE)

    def joint_array(self) -> np.ndarray:
        return np.array(self.joints, dtype=JOINT_DTYPE)

    @property
    def nb_bodies(self) -> int:
        return len(self.bodies)

    @property
    def nb_dynamic(self) -> int:
        return sum(1 for b in self.bodies if b[0] == DYNAMIC)


def make_settings(gravity=(0.0, -9.81, 0.0), vel_iters=6, pos_iters=3, friction=0.3, bounciness=0.5,
                  persistent=0.03, restitution_threshold=0.5, cos_similar=0.95, sleeping=0):
    s = np.zeros((), dtype=SETTINGS_DTYPE)
    s["gravity"] = gravity
    s["vel_iters"], s["pos_iters"] = vel_iters, pos_iters
    s["persistent"], s["restitution_threshold"], s["cos_similar"] = persistent, restitution_threshold, cos_similar
    s["default_friction"], s["default_bounciness"], s["sleeping"] = friction, bounciness, sleeping
    return s


# ---------------------------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------------------------

def box_pile(n_side=10, levels=10, spacing=1.5, jitter=0.2, seed=1, friction=0.5, bounciness=0.0):
    """1 m unit boxes (1 kg) on a grid dropped onto a static 100 x 100 x 1 m ground (gravity -z).
    n_side = levels = 10 is the 1000-box pile; smaller grids are parity cases."""
    rng = SplitMix64(seed)
    sc = Scene(f"box_pile_{n_side * n_side * levels}",
               make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5, friction=friction,
                             bounciness=bounciness))
    g = sc.add_body(STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, BOX, (50.0, 50.0, 0.5), friction=friction, bounciness=bounciness)
    half = (n_side - 1) / 2.0
    for k in range(levels):
        for j in range(n_side):
            for i in range(n_side):
                pos = ((i - half) * spacing + rng.uniform(-jitter, jitter),
                       (j - half) * spacing + rng.uniform(-jitter, jitter),
                       1.0 + spacing * k + rng.uniform(-jitter, jitter))
                b = sc.add_body(DYNAMIC, pos, rng.quaternion())
                sc.add_collider(b, BOX, (0.5, 0.5, 0.5), friction=friction, bounciness=bounciness, density=1.0)
    return sc


def mixed_container(n_bodies=65536, extent=(200.0, 200.0, 100.0), seed=2, vel_iters=6, pos_iters=3):
    """Spheres r~U[0.25,1], boxes with half-extents ~U[0.25,1], capsules r~U[0.2,0.5] h~U[0.5,2]
    (one third each, interleaved), density 1000, uniform positions in the volume, random
    orientations, v ~ N(0,1) per axis, inside a static floor + 4 walls of thickness 1 m."""
    rng = SplitMix64(seed)
    sc = Scene(f"mixed_container_{n_bodies}",
               make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=vel_iters, pos_iters=pos_iters))
    ex, ey, ez = extent
    hx, hy = ex / 2.0, ey / 2.0
    walls = [((0.0, 0.0, -0.5), (hx + 1.0, hy + 1.0, 0.5)),
             ((hx + 0.5, 0.0, ez / 2.0), (0.5, hy + 1.0, ez / 2.0)),
             ((-hx - 0.5, 0.0, ez / 2.0), (0.5, hy + 1.0, ez / 2.0)),
             ((0.0, hy + 0.5, ez / 2.0), (hx, 0.5, ez / 2.0)),
             ((0.0, -hy - 0.5, ez / 2.0), (hx, 0.5, ez / 2.0))]
    for pos, he in walls:
        w = sc.add_body(STATIC, pos)
        sc.add_collider(w, BOX, he)
    for i in range(n_bodies):
        pos = (rng.uniform(-hx, hx), rng.uniform(-hy, hy), rng.uniform(0.0, ez))
        q = rng.quaternion()
        kind = i % 3
        if kind == 0:
            params = (rng.uniform(0.25, 1.0),)
            shape = SPHERE
        elif kind == 1:
            params = (rng.uniform(0.25, 1.0), rng.uniform(0.25, 1.0), rng.uniform(0.25, 1.0))
            shape = BOX
        else:
            params = (rng.uniform(0.2, 0.5), rng.uniform(0.5, 2.0))
            shape = CAPSULE
        lin = (rng.normal(), rng.normal(), rng.normal())
        b = sc.add_body(DYNAMIC, pos, q, lin=lin)
        sc.add_collider(b, shape, params, density=1000.0)
    return sc


def trigger_scene(n_bodies=60, seed=12, extent=(10.0, 10.0, 12.0)):
    """Trigger colliders (Collider::setIsTrigger) among simulated bodies: static trigger zones (a box, a
    sphere, a capsule and a convex mesh) above a static ground, dynamic spheres / boxes / capsules /
    hulls falling through them, bodies whose only collider is a trigger (they fall through everything,
    reporting overlaps), compound bodies with a simulation box and a trigger sphere, and one collider that
    is neither a trigger nor a simulation collider (never tested)."""
    rng = SplitMix64(seed)
    sc = Scene(f"triggers_{n_bodies}", make_settings(gravity=(0.0, 0.0, -9.81)))
    ex, ey, ez = extent
    g = sc.add_body(STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, BOX, (ex, ey, 0.5))
    zones = sc.add_body(STATIC, (0.0, 0.0, 0.0))
    sc.add_collider(zones, BOX, (ex / 2, ey / 2, 0.6), lpos=(0.0, 0.0, 3.0), trigger=True)
    sc.add_collider(zones, SPHERE, (2.0,), lpos=(2.0, 2.0, 6.5), trigger=True)
    sc.add_collider(zones, CAPSULE, (0.8, 6.0), lpos=(-2.0, -1.0, 5.0), lquat=(0.7071068, 0.0, 0.0, 0.7071068),
                    trigger=True)
    h = sc.add_hull_from_points(ellipsoid_points(rng, 24, 1.0, 2.0))
    sc.add_collider(zones, HULL, (float(h),), lpos=(-2.5, 2.5, 8.0), trigger=True)
    hull_small = sc.add_hull_from_points(ellipsoid_points(rng, 24, 0.3, 0.6))
    for i in range(n_bodies):
        pos = (rng.uniform(-ex / 2, ex / 2), rng.uniform(-ey / 2, ey / 2), rng.uniform(2.0, ez))
        b = sc.add_body(DYNAMIC, pos, rng.quaternion(), lin=(0.3 * rng.normal(), 0.3 * rng.normal(), 0.0))
        kind = i % 6
        if kind == 0:
            sc.add_collider(b, SPHERE, (rng.uniform(0.25, 0.6),))
        elif kind == 1:
            sc.add_collider(b, BOX, (rng.uniform(0.25, 0.6), rng.uniform(0.25, 0.6), rng.uniform(0.25, 0.6)))
        elif kind == 2:
            sc.add_collider(b, CAPSULE, (rng.uniform(0.2, 0.4), rng.uniform(0.5, 1.2)))
        elif kind == 3:
            sc.add_collider(b, HULL, (float(hull_small),), density=500.0)
        elif kind == 4:   # only a trigger: falls through everything
            shape = (SPHERE, CAPSULE, BOX)[i % 3]
            params = {SPHERE: (0.4,), CAPSULE: (0.3, 0.8), BOX: (0.3, 0.3, 0.3)}[shape]
            sc.add_collider(b, shape, params, trigger=True)
        else:             # compound: a simulation box and a trigger sphere around it
            sc.add_collider(b, BOX, (0.35, 0.35, 0.35))
            sc.add_collider(b, SPHERE, (0.7,), trigger=True)
            if i % 12 == 11:   # and a third collider that is neither (never tested)
                sc.add_collider(b, SPHERE, (0.9,), lpos=(0.0, 0.0, 0.2), simulation=False)
    return sc


def sphere_chains(n_chains=16, length=16, radius=0.3, spacing=3.0, seed=3, settings=None, sc=None):
    """Chains of `length` spheres linked by ball-and-socket joints at their touching points; the first
    sphere of each chain hangs from a static anchor body (no collider).  Chains start horizontal in a
    seeded direction so that they swing down under gravity -z."""
    rng = SplitMix64(seed)
    if sc is None:
        sc = Scene(f"sphere_chains_{n_chains}x{length}",
                   settings if settings is not None else make_settings(gravity=(0.0, 0.0, -9.81)))
    side = int(math.ceil(math.sqrt(n_chains)))
    for c in range(n_chains):
        ax = (c % side - side / 2.0) * spacing * length * radius
        ay = (c // side - side / 2.0) * spacing * length * radius
        az = 2.0 * length * radius + 5.0
        ang = 2.0 * math.pi * rng.uniform01()
        d = (math.cos(ang), math.sin(ang), 0.0)
        anchor = sc.add_body(STATIC, (ax, ay, az))
        prev = anchor
        for k in range(length):
            p = (ax + d[0] * radius * (2 * k + 1), ay + d[1] * radius * (2 * k + 1), az)
            b = sc.add_body(DYNAMIC, p, rng.quaternion())
            sc.add_collider(b, SPHERE, (radius,))
            joint_at = (ax + d[0] * radius * 2 * k, ay + d[1] * radius * 2 * k, az)
            sc.add_joint(BALLSOCKET, prev, b, joint_at)
            prev = b
    return sc


def jointed_tower(height=6, seed=4, kinds=(HINGE, FIXED), friction=0.5, gravity=(0.0, 0.0, -9.81), sc=None,
                  origin=(0.0, 0.0), with_ground=True, limits=None):
    """A tower of 1 m boxes on a static ground; consecutive boxes are linked by joints (cycling through
    `kinds`) anchored at their shared face, hinge / slider axes seeded in the horizontal plane.
    `limits` maps a joint kind to the (flags, limits[4]) its joints get (none by default)."""
    rng = SplitMix64(seed)
    if sc is None:
        sc = Scene(f"jointed_tower_{height}", make_settings(gravity=gravity, vel_iters=10, pos_iters=5,
                                                             friction=friction, bounciness=0.0))
    if with_ground:
        g = sc.add_body(STATIC, (origin[0], origin[1], -0.5))
        sc.add_collider(g, BOX, (20.0, 20.0, 0.5), friction=friction, bounciness=0.0)
    prev = None
    for k in range(height):
        pos = (origin[0] + rng.uniform(-0.05, 0.05), origin[1] + rng.uniform(-0.05, 0.05), 0.5 + 1.0 * k)
        b = sc.add_body(DYNAMIC, pos)
        sc.add_collider(b, BOX, (0.5, 0.5, 0.5), friction=friction, bounciness=0.0)
        if prev is not None:
            kind = kinds[(k - 1) % len(kinds)]
            anchor = (pos[0], pos[1], pos[2] - 0.5)
            a = 2.0 * math.pi * rng.uniform01()
            fl, lim = (limits or {}).get(kind, (0, (0.0, 0.0, 0.0, 0.0)))
            sc.add_joint(kind, prev, b, anchor, axis=(math.cos(a), math.sin(a), 0.0), flags=fl, limits=lim)
        prev = b
    return sc


def cone_chains(n_chains=6, length=6, radius=0.3, half_angle=(0.2, 0.6), spacing=3.0, seed=7, sc=None):
    """Sphere chains like sphere_chains whose ball-and-socket joints carry a cone limit with a half angle
    ~U[half_angle].  The static anchor sits one radius behind the first joint, so that both cone axes
    (anchor minus centre of mass, for each body) of every joint lie along the chain at rest; spheres
    start with angular velocities ~N(0, 2) per axis so that the swinging chains bend past their cones."""
    rng = SplitMix64(seed)
    if sc is None:
        sc = Scene(f"cone_chains_{n_chains}x{length}", make_settings(gravity=(0.0, 0.0, -9.81)))
    side = int(math.ceil(math.sqrt(n_chains)))
    for c in range(n_chains):
        ax = (c % side - side / 2.0) * spacing * length * radius
        ay = (c // side - side / 2.0) * spacing * length * radius
        az = 2.0 * length * radius + 5.0
        ang = 2.0 * math.pi * rng.uniform01()
        d = (math.cos(ang), math.sin(ang), 0.0)
        prev = sc.add_body(STATIC, (ax - d[0] * radius, ay - d[1] * radius, az))
        for k in range(length):
            p = (ax + d[0] * radius * (2 * k + 1), ay + d[1] * radius * (2 * k + 1), az)
            w = (2.0 * rng.normal(), 2.0 * rng.normal(), 2.0 * rng.normal())
            b = sc.add_body(DYNAMIC, p, rng.quaternion(), ang=w)
            sc.add_collider(b, SPHERE, (radius,))
            joint_at = (ax + d[0] * radius * 2 * k, ay + d[1] * radius * 2 * k, az)
            cone = rng.uniform(half_angle[0], half_angle[1])
            sc.add_joint(BALLSOCKET, prev, b, joint_at, flags=1, limits=(cone, 0.0, 0.0, 0.0))
            prev = b
    return sc


def slider_rigs(n_rigs=8, spacing=4.0, seed=8, sc=None):
    """Two-stage slider rigs without colliders: a static post, a carriage box (half extents 0.3 m,
    density 1) 0.8 m below it on a slider whose axis is tilted 20-70 degrees from the vertical, up or
    down by the seed (limits [-0.4, 0.6] m), and a second box 0.8 m beside the carriage on a slider
    along a seeded horizontal axis (limits [-0.3, 0.3] m).  Rig r has a motor (speed ~U[-1, 1] m/s,
    max force 1 N) on the carriage when r % 3 == 1, on the second box when r % 3 == 2.  Gravity
    -z drives the carriages into one of their limits, the second box swings it around."""
    rng = SplitMix64(seed)
    if sc is None:
        sc = Scene(f"slider_rigs_{n_rigs}", make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5))
    side = int(math.ceil(math.sqrt(n_rigs)))
    for r in range(n_rigs):
        px, py, pz = (r % side) * spacing, (r // side) * spacing, 5.0
        post = sc.add_body(STATIC, (px, py, pz))
        tilt = math.radians(rng.uniform(20.0, 70.0))
        yaw = 2.0 * math.pi * rng.uniform01()
        sgn = 1.0 if rng.uniform01() < 0.5 else -1.0
        axis1 = (math.sin(tilt) * math.cos(yaw), math.sin(tilt) * math.sin(yaw), sgn * math.cos(tilt))
        c = (px, py, pz - 0.8)
        car = sc.add_body(DYNAMIC, c, rng.quaternion())
        sc.add_collider(car, BOX, (0.3, 0.3, 0.3))
        m1 = r % 3 == 1
        sc.add_joint(SLIDER, post, car, c, axis=axis1, flags=1 | (2 if m1 else 0),
                     limits=(-0.4, 0.6, rng.uniform(-1.0, 1.0) if m1 else 0.0, 1.0 if m1 else 0.0))
        yaw2 = 2.0 * math.pi * rng.uniform01()
        p2 = (c[0] + 0.8 * math.cos(yaw2), c[1] + 0.8 * math.sin(yaw2), c[2])
        box = sc.add_body(DYNAMIC, p2, rng.quaternion())
        sc.add_collider(box, BOX, (0.3, 0.3, 0.3))
        a2 = 2.0 * math.pi * rng.uniform01()
        m2 = r % 3 == 2
        sc.add_joint(SLIDER, car, box, p2, axis=(math.cos(a2), math.sin(a2), 0.0), flags=1 | (2 if m2 else 0),
                     limits=(-0.3, 0.3, rng.uniform(-1.0, 1.0) if m2 else 0.0, 1.0 if m2 else 0.0))
    return sc


def hinge_rigs(n_rigs=8, spacing=5.0, seed=9, sc=None):
    """Two-paddle hinge rigs without static colliders: a static post and two paddles (boxes 1.0 x 0.2 x
    0.6 m, density 1) hinged end to end on a horizontal axis (seeded yaw), starting horizontal.  Paddle
    1 has limits [-0.6, 0.4] rad, paddle 2 limits [-0.5, 0.5] rad; rig r has a motor (speed ~U[-2, 2]
    rad/s, max torque 0.5 N m) on paddle 1 when r % 3 == 1, on paddle 2 when r % 3 == 2.  Gravity -z
    swings the paddles into their limits."""
    rng = SplitMix64(seed)
    if sc is None:
        sc = Scene(f"hinge_rigs_{n_rigs}", make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5))
    side = int(math.ceil(math.sqrt(n_rigs)))
    for r in range(n_rigs):
        px, py, pz = (r % side) * spacing, (r // side) * spacing, 5.0
        post = sc.add_body(STATIC, (px, py, pz))
        yaw = 2.0 * math.pi * rng.uniform01()
        d = (math.cos(yaw), math.sin(yaw), 0.0)
        axis = (-d[1], d[0], 0.0)
        q = (0.0, 0.0, math.sin(yaw / 2.0), math.cos(yaw / 2.0))
        prev, lims = post, ((-0.6, 0.4), (-0.5, 0.5))
        for k in range(2):
            centre = (px + d[0] * (0.6 + 1.1 * k), py + d[1] * (0.6 + 1.1 * k), pz)
            pad = sc.add_body(DYNAMIC, centre, q)
            sc.add_collider(pad, BOX, (0.5, 0.1, 0.3))
            hinge_at = (px + d[0] * (0.05 + 1.1 * k), py + d[1] * (0.05 + 1.1 * k), pz)
            motor = r % 3 == k + 1
            sc.add_joint(HINGE, prev, pad, hinge_at, axis=axis, flags=1 | (2 if motor else 0),
                         limits=(lims[k][0], lims[k][1], rng.uniform(-2.0, 2.0) if motor else 0.0,
                                 0.5 if motor else 0.0))
            prev = pad
    return sc


def convex_hull(points):
    """Convex hull of a point cloud as PhysicsCommon::createConvexMesh(VertexArray) builds it (QuickHull, in
    the library's host code): (hull vertices [V, 3] float32 numbered in order of first use, faces as CCW
    index lists in the reference's face order)."""
    from reactphysics3d_amd import lib
    v, counts, idx = lib.quickhull(points)
    faces, o = [], 0
    for c in counts:
        faces.append([int(i) for i in idx[o:o + c]])
        o += c
    return v, faces


def ellipsoid_points(rng, n_points=24, lo=0.3, hi=1.2):
    """n points on an ellipsoid with semi-axes ~U[lo, hi] (point = axes * uniform unit vector)."""
    axes = [rng.uniform(lo, hi) for _ in range(3)]
    pts = []
    for _ in range(n_points):
        z = 2.0 * rng.uniform01() - 1.0
        a = 2.0 * math.pi * rng.uniform01()
        r = math.sqrt(max(0.0, 1.0 - z * z))
        pts.append((axes[0] * r * math.cos(a), axes[1] * r * math.sin(a), axes[2] * z))
    return np.array(pts, np.float32)


def ellipsoid_hull(rng, n_points=24, lo=0.3, hi=1.2):
    """QuickHull of n points on an ellipsoid with semi-axes ~U[lo, hi]: (vertices, faces)."""
    return convex_hull(ellipsoid_points(rng, n_points, lo, hi))


def hull_scene(n_hulls=16384, n_chains=256, chain_length=16, extent=(120.0, 120.0, 80.0), seed=3, n_points=24):
    """Convex-mesh bodies (hulls of 24 seeded ellipsoid points, semi-axes U[0.3, 1.2]) at seeded
    positions / orientations in a volume, plus chains of 0.3 m spheres on ball-and-socket joints
    hanging from static anchors, above a static ground box."""
    rng = SplitMix64(seed)
    sc = Scene(f"hulls_{n_hulls}_chains_{n_chains}x{chain_length}", make_settings(gravity=(0.0, 0.0, -9.81)))
    g = sc.add_body(STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, BOX, (extent[0] / 2 + 10.0, extent[1] / 2 + 10.0, 0.5))
    for _ in range(n_hulls):
        h = sc.add_hull_from_points(ellipsoid_points(rng, n_points))
        pos = (rng.uniform(-extent[0] / 2, extent[0] / 2), rng.uniform(-extent[1] / 2, extent[1] / 2),
               rng.uniform(1.0, extent[2]))
        b = sc.add_body(DYNAMIC, pos, rng.quaternion())
        sc.add_collider(b, HULL, (float(h),), density=1000.0)
    if n_chains:
        sphere_chains(n_chains, chain_length, radius=0.3, spacing=3.0, seed=seed + 1, sc=sc)
    return sc


def terrain_mesh(nx=12, ny=12, cell=1.0, amp=0.35, seed=9):
    """A bumpy grid terrain as a triangle mesh in the xy plane (z up): (nx + 1) x (ny + 1) vertices
    with seeded heights ~U[-amp, amp], two counter-clockwise triangles per cell with the diagonal
    alternating from cell to cell."""
    rng = SplitMix64(seed)
    v = np.array([((i - nx / 2) * cell, (j - ny / 2) * cell, rng.uniform(-amp, amp))
                  for j in range(ny + 1) for i in range(nx + 1)], np.float32)
    t = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i
            if (i + j) % 2 == 0:
                t += [(a, b, c), (a, c, d)]
            else:
                t += [(a, b, d), (b, c, d)]
    return v, np.array(t, np.int32)


def concave_scene(kind="mesh", n_bodies=40, shapes=(SPHERE, CAPSULE, BOX, HULL), nx=10, scale=(1.0, 1.0, 1.0),
                  seed=9, height=(1.5, 6.0), normals=False, ground_type=STATIC, ground_lin=(0.0, 0.0, 0.0),
                  ground_ang=(0.0, 0.0, 0.0), sleeping=False, trigger_every=0):
    """Dynamic bodies of the given shapes (cycled) dropped at seeded positions and orientations onto a
    concave ground: a bumpy triangle mesh (kind "mesh", z up; normals=True hands the mesh its own
    vertex normals instead of letting the library compute them) or a height field (kind "heightfield",
    y up, gravity -y) of nx x nx cells, with the given shape scale.  The ground is static, or kinematic
    with the given velocities; sleeping turns the WorldSettings' sleeping on; trigger_every = n makes
    every n-th falling body's collider a trigger."""
    rng = SplitMix64(seed)
    up = 2 if kind == "mesh" else 1
    g = [0.0, 0.0, 0.0]
    g[up] = -9.81
    sc = Scene(f"concave_{kind}_{n_bodies}", make_settings(gravity=tuple(g), sleeping=1 if sleeping else 0))
    ground = sc.add_body(ground_type, (0.0, 0.0, 0.0), lin=ground_lin, ang=ground_ang)
    if kind == "mesh":
        v, t = terrain_mesh(nx, nx, seed=seed + 1)
        n = None
        if normals:
            n = np.zeros_like(v)
            n[:, 2] = 1.0
            n[:, 0] = 0.05 * np.sin(v[:, 0])
            n /= np.linalg.norm(n, axis=1, keepdims=True)
        m = sc.add_triangle_mesh(v, t, n)
        sc.add_collider(ground, CONCAVE_MESH, (float(m),) + tuple(scale))
    else:
        h = np.array([[rng.uniform(-0.35, 0.35) for _ in range(nx + 1)] for _ in range(nx + 1)], np.float32)
        f = sc.add_heightfield(h)
        sc.add_collider(ground, HEIGHTFIELD, (float(f),) + tuple(scale))
    half = 0.5 * nx * min(scale[0], scale[2] if kind != "mesh" else scale[1]) - 1.0
    for k in range(n_bodies):
        shape = shapes[k % len(shapes)]
        pos = [rng.uniform(-half, half), rng.uniform(-half, half), rng.uniform(-half, half)]
        pos[up] = rng.uniform(height[0], height[1])
        b = sc.add_body(DYNAMIC, tuple(pos), rng.quaternion())
        trig = trigger_every > 0 and k % trigger_every == trigger_every - 1
        if shape == SPHERE:
            sc.add_collider(b, SPHERE, (rng.uniform(0.2, 0.5),), trigger=trig)
        elif shape == CAPSULE:
            sc.add_collider(b, CAPSULE, (rng.uniform(0.15, 0.3), rng.uniform(0.3, 0.8)), trigger=trig)
        elif shape == BOX:
            sc.add_collider(b, BOX, (rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.5)), trigger=trig)
        else:
            hh = sc.add_hull_from_points(ellipsoid_points(rng, 16, 0.25, 0.5))
            sc.add_collider(b, HULL, (float(hh),), trigger=trig)
    return sc


def sleeping_pile(n_side=3, levels=2, seed=5, ball_height=None, ball_radius=0.4):
    """box_pile with sleeping enabled (WorldSettings defaults: 1 s still below 0.02 m/s and 3 deg/s);
    with ball_height, a dense sphere (density 5) dropped from that height near the pile's centre
    lands on the pile after it fell asleep and wakes the island it hits."""
    sc = box_pile(n_side, levels, seed=seed)
    sc.name = f"sleeping_pile_{n_side * n_side * levels}" + ("_ball" if ball_height else "")
    sc.settings["sleeping"] = 1
    if ball_height:
        b = sc.add_body(DYNAMIC, (0.3, 0.2, ball_height))
        sc.add_collider(b, SPHERE, (ball_radius,), density=5.0)
    return sc


def sleeping_chain(length=6, radius=0.3, hit_distance=5.0, hit_speed=3.0, seed=6):
    """A vertical chain of spheres hanging at rest from a static anchor by ball-and-socket joints, with
    sleeping enabled: it falls asleep after the time before sleep.  A sphere without gravity flies at
    hit_speed towards the lowest sphere from hit_distance away, hits it after the chain slept, and
    wakes the whole chain through its joints.  Friction 0.5 and no bounce, as in box_pile, so that both
    can share a batch."""
    rng = SplitMix64(seed)
    sc = Scene(f"sleeping_chain_{length}", make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5,
                                                         friction=0.5, bounciness=0.0, sleeping=1))
    top = 2.0 * radius * length + 2.0
    prev = sc.add_body(STATIC, (0.0, 0.0, top + radius))
    for k in range(length):
        z = top - radius * (2 * k + 1)
        b = sc.add_body(DYNAMIC, (0.0, 0.0, z), rng.quaternion())
        sc.add_collider(b, SPHERE, (radius,))
        sc.add_joint(BALLSOCKET, prev, b, (0.0, 0.0, z + radius))
        prev = b
    z_low = top - radius * (2 * length - 1)
    ball = sc.add_body(DYNAMIC, (hit_distance, 0.05, z_low), lin=(-hit_speed, 0.0, 0.0), gravity=False)
    sc.add_collider(ball, SPHERE, (radius,), density=2.0)
    return sc


def body_features(seed=10):
    """Bodies beyond plain dynamic ones: a kinematic platform (translating at 0.5 m/s and turning at
    0.3 rad/s, 0.25 m above the ground) carrying two boxes, spheres with linear damping 0.5 and angular
    damping 0.8, a box whose motion is locked to the x-z plane and to rotation about y (lock factors
    (1, 0, 1) and (0, 1, 0)), and a capsule without gravity drifting and spinning into the pile."""
    rng = SplitMix64(seed)
    sc = Scene("body_features", make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5, friction=0.5,
                                              bounciness=0.2))
    g = sc.add_body(STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, BOX, (20.0, 20.0, 0.5))
    plat = sc.add_body(KINEMATIC, (0.0, 0.0, 0.5), lin=(0.5, 0.0, 0.0), ang=(0.0, 0.0, 0.3))
    sc.add_collider(plat, BOX, (2.0, 2.0, 0.25))
    for k in range(2):
        b = sc.add_body(DYNAMIC, (-0.8 + 1.6 * k, 0.3 * k, 1.06), rng.quaternion() if k else (0.0, 0.0, 0.0, 1.0))
        sc.add_collider(b, BOX, (0.3, 0.3, 0.3))
    for k in range(3):
        b = sc.add_body(DYNAMIC, (4.0 + 0.9 * k, rng.uniform(-0.3, 0.3), 1.0 + 1.2 * k),
                        ang=(rng.normal(), rng.normal(), 4.0 * rng.normal()), ldamp=0.5, adamp=0.8)
        sc.add_collider(b, SPHERE, (0.4,))
    lk = sc.add_body(DYNAMIC, (-4.0, 0.0, 2.0), rng.quaternion(), lin=(1.0, 1.0, 0.0), ang=(2.0, 1.0, 3.0),
                     llock=(1.0, 0.0, 1.0), alock=(0.0, 1.0, 0.0))
    sc.add_collider(lk, BOX, (0.4, 0.3, 0.5))
    cap = sc.add_body(DYNAMIC, (8.0, 0.0, 0.9), rng.quaternion(), lin=(-1.5, 0.0, 0.0), ang=(0.0, 2.0, 0.0),
                      gravity=False)
    sc.add_collider(cap, CAPSULE, (0.3, 1.2))
    return sc


def batched_world(world_id, seed=4, min_bodies=8, max_bodies=512, max_joints=32):
    """One world of the many-world batch: n = round(exp(U[ln min, ln max])) unit boxes (1 kg, friction
    0.5, restitution 0) on a static ground, either a pile (seeded grid, jitter 0.2 m, random
    orientations, 1.5 m spacing) or stacks of height U{4..20}; U{0..max_joints} hinge / fixed joints
    between neighbouring boxes.  Stream seed = seed + world_id."""
    rng = SplitMix64(seed + world_id)
    lo, hi = math.log(min_bodies), math.log(max_bodies)
    n = int(round(math.exp(lo + (hi - lo) * rng.uniform01())))
    n = max(min_bodies, min(max_bodies, n))
    stack = rng.uniform01() < 0.5
    sc = Scene(f"batched_{world_id}", make_settings(gravity=(0.0, 0.0, -9.81), vel_iters=10, pos_iters=5,
                                                    friction=0.5, bounciness=0.0))
    g = sc.add_body(STATIC, (0.0, 0.0, -0.5))
    sc.add_collider(g, BOX, (50.0, 50.0, 0.5), friction=0.5, bounciness=0.0)
    cand = []   # neighbouring body pairs a joint may link, with the anchor between them
    if stack:
        h = 4 + min(16, int(17 * rng.uniform01()))
        cols = (n + h - 1) // h
        side = int(math.ceil(math.sqrt(cols)))
        k = 0
        for c in range(cols):
            cx, cy = (c % side - (side - 1) / 2.0) * 2.0, (c // side - (side - 1) / 2.0) * 2.0
            prev = None
            for lvl in range(h):
                if k == n:
                    break
                pos = (cx + rng.uniform(-0.05, 0.05), cy + rng.uniform(-0.05, 0.05), 0.5 + 1.01 * lvl)
                b = sc.add_body(DYNAMIC, pos)
                sc.add_collider(b, BOX, (0.5, 0.5, 0.5), friction=0.5, bounciness=0.0, density=1.0)
                if prev is not None:
                    cand.append((prev, b, (pos[0], pos[1], pos[2] - 0.505)))
                prev = b
                k += 1
    else:
        levels = 1 + min(7, int(8 * rng.uniform01()))
        side = int(math.ceil(math.sqrt(n / levels)))
        half = (side - 1) / 2.0
        k = 0
        grid = {}
        for lvl in range(levels + 1):
            for j in range(side):
                for i in range(side):
                    if k == n:
                        break
                    pos = ((i - half) * 1.5 + rng.uniform(-0.2, 0.2), (j - half) * 1.5 + rng.uniform(-0.2, 0.2),
                           1.0 + 1.5 * lvl + rng.uniform(-0.2, 0.2))
                    b = sc.add_body(DYNAMIC, pos, rng.quaternion())
                    sc.add_collider(b, BOX, (0.5, 0.5, 0.5), friction=0.5, bounciness=0.0, density=1.0)
                    grid[(i, j, lvl)] = (b, pos)
                    if (i - 1, j, lvl) in grid:
                        a, pa = grid[(i - 1, j, lvl)]
                        cand.append((a, b, tuple((pa[t] + pos[t]) * 0.5 for t in range(3))))
                    k += 1
    nj = min(len(cand), int((max_joints + 1) * rng.uniform01()))
    for t in range(nj):   # partial Fisher-Yates over the candidates
        r = t + min(len(cand) - 1 - t, int((len(cand) - t) * rng.uniform01()))
        cand[t], cand[r] = cand[r], cand[t]
        a, b, anchor = cand[t]
        kind = HINGE if rng.uniform01() < 0.5 else FIXED
        ang = 2.0 * math.pi * rng.uniform01()
        sc.add_joint(kind, a, b, anchor, axis=(math.cos(ang), math.sin(ang), 0.0))
    return sc


def batched_worlds(n_worlds=4096, seed=4, **kw):
    """The many-world batch: world w is batched_world(w, seed)."""
    return [batched_world(w, seed, **kw) for w in range(n_worlds)]


def two_body(shape1, params1, shape2, params2, pos1, pos2, q1=(0, 0, 0, 1), q2=(0, 0, 0, 1), static2=False):
    """Two bodies, no gravity: the unit case of each narrowphase pair type."""
    sc = Scene("two_body", make_settings(gravity=(0.0, 0.0, 0.0)))
    a = sc.add_body(DYNAMIC, pos1, q1)
    sc.add_collider(a, shape1, params1)
    b = sc.add_body(STATIC if static2 else DYNAMIC, pos2, q2)
    sc.add_collider(b, shape2, params2)
    return sc
