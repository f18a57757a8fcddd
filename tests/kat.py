"""Known-answer narrowphase cases from the reference's own tests (tests/golden/world_queries_kat.json),
turned into two-body scenes (no gravity, no motion) and checked against contacts reported by a step."""
import json
import math
import os

import numpy as np

from reactphysics3d_amd import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "golden", "world_queries_kat.json")) as f:
        return json.load(f)


def euler_quat(ax, ay, az):
    """Quaternion from Euler angles, the convention of Quaternion::initWithEulerAngles
    (src/mathematics/Quaternion.cpp:207-233), normalised."""
    sx, cx = math.sin(ax * 0.5), math.cos(ax * 0.5)
    sy, cy = math.sin(ay * 0.5), math.cos(ay * 0.5)
    sz, cz = math.sin(az * 0.5), math.cos(az * 0.5)
    q = np.array([sx * cy * cz - cx * sy * sz, cx * sy * cz + sx * cy * sz, cx * cy * sz - sx * sy * cz,
                  cx * cy * cz + sx * sy * sz])
    return q / np.linalg.norm(q)


# the fixture's convex meshes (TestWorldQueries.h:377-434): a box given as 8 vertices and 6 quads
_CUBE_SIGNS = [(-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1), (-1, 1, -1)]
_CUBE_FACES = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3]]


def _shape(sc, spec):
    kind, params = spec
    if kind == "convex":
        v = np.array([[sx * params[0], sy * params[1], sz * params[2]] for sx, sy, sz in _CUBE_SIGNS], np.float32)
        return S.HULL, (float(sc.add_hull(v, _CUBE_FACES)),)
    return {"sphere": S.SPHERE, "capsule": S.CAPSULE, "box": S.BOX}[kind], tuple(params)


def scene(case):
    sc = S.Scene(case["name"], S.make_settings(gravity=(0.0, 0.0, 0.0)))
    q1 = tuple(euler_quat(*case["euler1"]).astype(np.float32))
    q2 = tuple(euler_quat(*case["euler2"]).astype(np.float32))
    s1 = _shape(sc, case["shape1"])
    s2 = _shape(sc, case["shape2"])
    b1 = sc.add_body(S.DYNAMIC, case["pos1"], q1)
    sc.add_collider(b1, *s1)
    b2 = sc.add_body(S.STATIC, case["pos2"], q2)
    sc.add_collider(b2, *s2)
    return sc, q2


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def check(case, contacts, eps):
    """contacts: dict of arrays as returned by Batch.contacts() / RefWorld.contacts().  Returns a list of
    problems (empty when every expected point is matched within eps, either body order)."""
    _, q2 = scene(case)
    R2 = _rot(np.array(q2, np.float64))
    exp = []
    for p in case["points"]:
        l2 = p.get("local2")
        if l2 is None:
            l2 = R2.T @ (np.array(p["world2"], np.float64) - np.array(case["pos2"], np.float64))
        exp.append((np.array(p["local1"], np.float64), np.array(l2, np.float64), p["depth"]))
    if len(contacts["colliders"]) != 1:
        return [f"{case['name']}: {len(contacts['colliders'])} manifolds, expected 1"]
    n = int(contacts["nb_points"][0])
    if n != len(exp):
        return [f"{case['name']}: {n} contact points, expected {len(exp)}"]
    swapped = int(contacts["colliders"][0][0]) != 0
    got = []
    for k in range(n):
        a, b = contacts["local1"][0][k].astype(np.float64), contacts["local2"][0][k].astype(np.float64)
        got.append((b, a) if swapped else (a, b))
    problems = []
    for l1, l2, d in exp:
        ok = any(np.all(np.abs(g1 - l1) <= eps) and np.all(np.abs(g2 - l2) <= eps) and
                 abs(float(contacts["normal_depth"][0][k][3]) - d) <= eps for k, (g1, g2) in enumerate(got))
        if not ok:
            problems.append(f"{case['name']}: no contact point near {l1}, {l2}, depth {d}; got {got}")
    return problems
