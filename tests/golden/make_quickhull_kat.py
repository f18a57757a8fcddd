"""Writes quickhull_kat.json: the five point clouds of the reference's QuickHull test
(test/tests/utils/TestQuickHull.h:70-230) and, for each, the indices of the input points the
reference's test requires among the hull's vertices.  Reads the reference test file as text (point
literals and indices only); run from the repository root with the reference at /root/reference."""
import json
import os
import re

SRC = "/root/reference/test/tests/utils/TestQuickHull.h"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    text = open(SRC).read()
    body = text[text.index("void testConvexHulls()"):text.index("bool testPointsAmongHullVertices")]
    cases = []
    for chunk in body.split("createConvexMesh(")[:-1]:
        pts = re.findall(r"points\.push_back\(Vector3\(([-\d.]+),\s*([-\d.]+),\s*([-\d.]+)\)\)", chunk)
        if pts:
            cases.append({"points": [[float(x) for x in p] for p in pts]})
    idx_chunks = body.split("indicesOfHull")
    groups, cur = [], []
    for line in body.splitlines():
        m = re.search(r"indicesOfHull\.push_back\((\d+)\)", line)
        if m:
            cur.append(int(m.group(1)))
        elif "testPointsAmongHullVertices" in line and cur:
            groups.append(cur)
            cur = []
    assert len(groups) == len(cases) == 5, (len(groups), len(cases))
    for c, g in zip(cases, groups):
        c["hull_indices"] = g
    with open(os.path.join(HERE, "quickhull_kat.json"), "w") as f:
        json.dump({"source": "test/tests/utils/TestQuickHull.h (ReactPhysics3D 0.10.2)", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
