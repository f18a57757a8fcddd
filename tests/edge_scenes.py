"""Two-body scenes for the degenerate branches of the narrowphase (no gravity unless stated):
coincident sphere centres (SphereVsSphereAlgorithm.cpp:92-99), parallel overlapping capsules
(CapsuleVsCapsuleAlgorithm.cpp:86-162, two clipped points, also with coincident inner segments),
a capsule lying on a convex-mesh face (the GJK shallow path and the SAT face path,
CapsuleVsConvexPolyhedronAlgorithm.cpp:80-144, SATAlgorithm.cpp:275-287: two points) and a
ball-and-socket cone limit whose anchor is the centre of mass of its body."""
import numpy as np

from reactphysics3d_amd import scene as S
from tests import kat


def coincident_spheres():
    sc = S.Scene("coincident_spheres", S.make_settings(gravity=(0.0, 0.0, 0.0)))
    a = sc.add_body(S.DYNAMIC, (1.0, 2.0, 3.0), (0.1, 0.2, 0.3, 0.9273618))
    sc.add_collider(a, S.SPHERE, (0.5,))
    b = sc.add_body(S.STATIC, (1.0, 2.0, 3.0))
    sc.add_collider(b, S.SPHERE, (0.8,))
    return sc


def parallel_capsules(gap=0.3, offset=0.6, coincident=False):
    """Two capsules along y (height 2, radii 0.5 / 0.4), side by side at distance `gap` between axes (0:
    the inner segments overlap), shifted by `offset` along the axis."""
    sc = S.Scene("parallel_capsules", S.make_settings(gravity=(0.0, 0.0, 0.0)))
    a = sc.add_body(S.DYNAMIC, (0.0, 0.0, 0.0))
    sc.add_collider(a, S.CAPSULE, (0.5, 2.0))
    b = sc.add_body(S.STATIC, (0.0 if coincident else gap, offset, 0.0))
    sc.add_collider(b, S.CAPSULE, (0.4, 2.0))
    return sc


def capsule_on_mesh_face(depth):
    """A capsule (radius 0.3, height 1.6) lying along x on the top face (z = 0.5) of a 2 x 2 x 1 convex-mesh
    box, its surface `depth` below the face (small: the GJK margin case; large: the SAT face case)."""
    sc = S.Scene("capsule_on_face", S.make_settings(gravity=(0.0, 0.0, 0.0)))
    v = np.array([[sx * 1.0, sy * 1.0, sz * 0.5] for sx, sy, sz in kat._CUBE_SIGNS], np.float32)
    h = sc.add_hull(v, kat._CUBE_FACES)
    box = sc.add_body(S.STATIC, (0.0, 0.0, 0.0))
    sc.add_collider(box, S.HULL, (float(h),))
    # capsule axis is local y: rotate it onto x (quarter turn about z)
    c = sc.add_body(S.DYNAMIC, (0.1, 0.05, 0.5 + 0.3 - depth), (0.0, 0.0, 0.70710678, 0.70710678))
    sc.add_collider(c, S.CAPSULE, (0.3, 1.6))
    return sc


def cone_anchor_at_com():
    """A box hanging on a ball-and-socket joint from a static anchor body, the joint anchor at the box's
    centre of mass, with a cone limit; gravity on, an initial spin."""
    sc = S.Scene("cone_at_com", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    a = sc.add_body(S.STATIC, (0.0, 0.0, 5.0))
    sc.add_collider(a, S.SPHERE, (0.2,))
    b = sc.add_body(S.DYNAMIC, (0.0, 0.0, 4.0), ang=(1.5, -0.7, 0.4))
    sc.add_collider(b, S.BOX, (0.3, 0.2, 0.6))
    sc.add_joint(S.BALLSOCKET, a, b, (0.0, 0.0, 4.0), flags=1, limits=(0.4, 0.0, 0.0, 0.0))
    return sc


def prism_hull(n, radius=1.0, half_height=0.5):
    """An n-gon prism as a polygon mesh: bottom ring then top ring of vertices, two n-vertex caps and n
    quads, counter-clockwise seen from outside."""
    a = 2.0 * np.pi * np.arange(n) / n
    ring = np.stack([radius * np.cos(a), radius * np.sin(a)], 1)
    v = np.concatenate([np.c_[ring, -half_height * np.ones(n)], np.c_[ring, half_height * np.ones(n)]])
    faces = [list(range(n - 1, -1, -1)), list(range(n, 2 * n))]
    faces += [[i, (i + 1) % n, n + (i + 1) % n, n + i] for i in range(n)]
    return v.astype(np.float32), faces


def stacked_prisms(n=24, twist=0.07):
    """A 24-gon prism resting cap to cap on a static one, turned by `twist` radians about the common
    axis, gravity on: the face contact clips a 24-vertex incident cap by the 24 side planes of the
    reference cap, the largest polygon the clip buffers hold (48 vertices)."""
    sc = S.Scene("stacked_prisms", S.make_settings(gravity=(0.0, 0.0, -9.81)))
    v, f = prism_hull(n)
    h = sc.add_hull(v, f)
    base = sc.add_body(S.STATIC, (0.0, 0.0, 0.0))
    sc.add_collider(base, S.HULL, (float(h),))
    s, c = np.sin(twist / 2.0), np.cos(twist / 2.0)
    top = sc.add_body(S.DYNAMIC, (0.02, -0.01, 0.995), (0.0, 0.0, s, c))
    sc.add_collider(top, S.HULL, (float(h),))
    return sc
