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
