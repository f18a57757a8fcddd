"""Writes world_queries_kat.json: the known-answer contact points of the reference's narrowphase test
(test/tests/collision/TestWorldQueries.h) for the ten convex pair types, every sub-case (face, edge and
corner contacts, rotated bodies).

For each testXVsYCollision() function (concave ones excluded) every place where the test moves its two
bodies (setTransform(transform1), setTransform(transform2)) starts a case: the two transforms, the shapes
of the two bodies (from the fixture constructor, :325-437), the number of contact points the test
expects and the expected points (local point on body 1, local point on body 2, penetration depth).  The
test file is read as text and the C++ expressions of those values are evaluated in Python (sqrt, pi,
Quaternion::fromEulerAngles angles); run from the repository root with the reference at /root/reference."""
import json
import math
import os
import re

SRC = "/root/reference/test/tests/collision/TestWorldQueries.h"
HERE = os.path.dirname(os.path.abspath(__file__))

# bodies of the fixture constructor and their shapes
BODIES = {
    "mSphereBody1": ["sphere", [3.0]], "mSphereBody2": ["sphere", [5.0]],
    "mBoxBody1": ["box", [3.0, 3.0, 3.0]], "mBoxBody2": ["box", [4.0, 2.0, 8.0]],
    "mCapsuleBody1": ["capsule", [2.0, 6.0]], "mCapsuleBody2": ["capsule", [3.0, 4.0]],
    "mConvexMeshBody1": ["convex", [3.0, 3.0, 3.0]], "mConvexMeshBody2": ["convex", [4.0, 2.0, 8.0]],
}


class V(tuple):
    def __new__(cls, x, y, z):
        return super().__new__(cls, (float(x), float(y), float(z)))

    def __rmul__(self, s):
        return V(s * self[0], s * self[1], s * self[2])

    __mul__ = __rmul__


def _euler(a, b, c):
    return ("euler", float(a), float(b), float(c))


def _py(expr):
    e = expr.strip().rstrip(";")
    e = re.sub(r"(\d+\.\d*|\d*\.\d+|\d+)f\b", r"\1", e)
    e = e.replace("std::sqrt", "math.sqrt").replace("rp3d::PI_RP3D", "math.pi").replace("PI_RP3D", "math.pi")
    e = e.replace("Quaternion::identity()", "_euler(0, 0, 0)").replace("Quaternion::fromEulerAngles", "_euler")
    e = e.replace("Vector3", "V").replace("decimal(", "float(")
    return eval(e, {"math": math, "V": V, "_euler": _euler})


def _transform(args):
    # "Vector3(...), Quaternion..." -> (position, euler angles)
    depth, cut = 0, None
    for i, ch in enumerate(args):
        depth += ch == "("
        depth -= ch == ")"
        if ch == "," and depth == 0:
            cut = i
            break
    pos = _py(args[:cut])
    q = _py(args[cut + 1:])
    return list(pos), list(q[1:])


def main():
    text = open(SRC).read()
    funcs = re.split(r"\n\s*void (test\w+Collision)\(\) \{", text)
    cases = []
    for name, body in zip(funcs[1::2], funcs[2::2]):
        if "Concave" in name:
            continue
        t = {}
        bodies = [None, None]
        vars_ = {}
        cur = None
        for line in body.splitlines():
            s = line.strip()
            m = re.match(r"(?:Transform )?(transform[12])\s*(?:=\s*Transform)?\((.*)\);$", s)
            if m:
                t[m.group(1)] = _transform(m.group(2))
                continue
            m = re.match(r"(m\w+Body\d)->setTransform\((transform[12])\);", s)
            if m:
                k = 0 if m.group(2) == "transform1" else 1
                bodies[k] = m.group(1)
                if k == 1:   # both bodies placed: a new case
                    cur = {"name": f"{name[4:-9]}_{sum(c['name'].startswith(name[4:-9]) for c in cases)}",
                           "shape1": BODIES[bodies[0]], "shape2": BODIES[bodies[1]],
                           "pos1": t["transform1"][0], "euler1": t["transform1"][1],
                           "pos2": t["transform2"][0], "euler2": t["transform2"][1], "nb_points": None,
                           "points": []}
                    cases.append(cur)
                    vars_ = {}
                continue
            if cur is None:
                continue
            m = re.search(r"getTotalNbContactPoints\(\) == (\d+)", s)
            if m and cur["nb_points"] is None:
                cur["nb_points"] = int(m.group(1))
                continue
            m = re.match(r"(?:Vector3 |decimal )?(localBody[12]Point\d*|penetrationDepth\d*)\s*(?:=\s*(.*);|\((.*)\);)$", s)
            if m:
                expr = m.group(2) if m.group(2) is not None else f"Vector3({m.group(3)})" \
                    if m.group(1).startswith("local") else m.group(3)
                if "transform2.getInverse() *" in expr:   # a world point on body 2
                    vars_[m.group(1)] = ("world", _py(expr.split("*", 1)[1]))
                else:
                    vars_[m.group(1)] = _py(expr)
                continue
            m = re.search(r"hasContactPointSimilarTo\(swappedBodiesCollisionData \? (localBody2Point\d*)", s)
            if m:
                suffix = m.group(1)[len("localBody2Point"):]
                l2 = vars_["localBody2Point" + suffix]
                p = {"local1": list(vars_["localBody1Point" + suffix]), "depth": float(vars_["penetrationDepth" + suffix])}
                if isinstance(l2, tuple) and len(l2) == 2 and l2[0] == "world":
                    p["world2"] = list(l2[1])
                else:
                    p["local2"] = list(l2)
                if p not in cur["points"]:
                    cur["points"].append(p)
    cases = [c for c in cases if c["points"] and c["nb_points"] == len(c["points"])]
    out = {"source": "test/tests/collision/TestWorldQueries.h (ReactPhysics3D 0.10.2): every sub-case of the ten "
                     "convex pair tests (testSphereVsSphereCollision ... testCapsuleVsCapsuleCollision); shapes from "
                     "the fixture constructor :325-437.  Known-answer contact points (local point on body 1, local "
                     "point on body 2, penetration depth); the reference test accepts a point within 0.001.",
           "epsilon": 0.001, "cases": cases}
    with open(os.path.join(HERE, "world_queries_kat.json"), "w") as f:
        json.dump(out, f, indent=1)
    kinds = sorted({(c["shape1"][0], c["shape2"][0]) for c in cases})
    print(len(cases), "cases,", len(kinds), "pair types:", kinds)


if __name__ == "__main__":
    main()
