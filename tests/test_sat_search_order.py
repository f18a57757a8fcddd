"""Host check of the order argument behind the shared polyhedron search (narrowphase.hip, EdgeSearch).

The reference walks the edge pairs of two polyhedra in order (SATAlgorithm.cpp:749-815): it stops at
the first separating pair, and otherwise an edge pair replaces the face result when
pen * 1.002 + 0.0005 < the face penetration, later pairs replacing it when strictly smaller.  The GPU
kernel instead lets the 64 lanes of a wavefront take every 64th pair, keeps (first separating pair,
first pair of the smallest positive penetration) per lane, combines the lanes with a butterfly, and
applies the face test once to the combined minimum.  The two must choose the same pair for every
sequence, including ties, NaN, kDecimalLargest and values that round together under the face test.
The face searches (sat_faces, SATAlgorithm.cpp:1119-1143) reduce the same way from kDecimalLargest.
"""
import numpy as np

F32 = np.float32
REL, ABS = F32(1.002), F32(0.0005)
LARGEST = np.finfo(np.float32).max
INT_MAX = 2**31 - 1


def face_test(pen, min_pen):
    with np.errstate(over="ignore"):   # kDecimalLargest * 1.002 rounds to +inf, as in float32 on the GPU
        return F32(F32(pen * REL) + ABS) < min_pen


def walk_edges(valid, pens, min_pen0):
    """The reference's loop, pair by pair."""
    is_face, min_pen, best = True, F32(min_pen0), None
    for k, (v, p) in enumerate(zip(valid, pens)):
        if not v:
            continue
        if p <= 0:
            return ("sep", k)
        if (is_face and face_test(p, min_pen)) or (not is_face and p < min_pen):
            min_pen, is_face, best = p, False, k
    return ("face",) if is_face else ("edge", best)


def lane_visit(state, k, v, p):
    sep, kk, pen = state
    if not v:
        return state
    if p <= 0:
        return (min(sep, k), kk, pen)
    if p < pen:
        return (sep, k, p)
    return state


def combine(a, b):
    sep = min(a[0], b[0])
    if b[2] < a[2] or (b[2] == a[2] and b[1] < a[1]):
        return (sep, b[1], b[2])
    return (sep, a[1], a[2])


def shared_search(valid, pens, start, lanes=64):
    states = []
    for lane in range(lanes):
        st = (INT_MAX, INT_MAX, F32(start))
        for k in range(lane, len(pens), lanes):
            st = lane_visit(st, k, valid[k], pens[k])
        states.append(st)
    off = lanes // 2
    while off:   # butterfly: every lane ends with the same combined state
        states = [combine(states[i], states[i ^ off]) for i in range(lanes)]
        off //= 2
    assert all(s == states[0] or (s[0] == states[0][0] and s[1] == states[0][1]) for s in states)
    return states[0]


def shared_edges(valid, pens, min_pen0):
    sep, k, pen = shared_search(valid, pens, np.inf)
    if sep != INT_MAX:
        return ("sep", sep)
    if k != INT_MAX and face_test(pen, F32(min_pen0)):
        return ("edge", k)
    return ("face",)


def walk_faces(pens):
    mn, face = LARGEST, 0
    for f, p in enumerate(pens):
        if p <= 0:
            return p, f
        if p < mn:
            mn, face = p, f
    return mn, face


def shared_faces(pens):
    sep, k, pen = shared_search([True] * len(pens), pens, LARGEST)
    if sep != INT_MAX:
        return pens[sep], sep
    if k == INT_MAX:
        return LARGEST, 0
    return pen, k


def _sequence(rng, n):
    base = rng.uniform(0.001, 0.2, size=6).astype(np.float32)
    pool = np.concatenate([base, np.nextafter(base, np.float32(1)), [LARGEST, np.float32(np.nan),
                                                                     np.float32(-0.01), np.float32(0.0)]])
    p = rng.choice(np.concatenate([pool[:12]] * 6 + [pool[12:]]), size=n).astype(np.float32)
    if rng.random() < 0.6:   # most walks have no separating pair
        p = np.where(p <= 0, np.float32(0.05), p).astype(np.float32)
    valid = rng.random(n) < rng.uniform(0.2, 1.0)
    return valid.tolist(), [F32(x) for x in p]


def test_shared_edge_search_matches_the_walk():
    rng = np.random.default_rng(7)
    kinds = set()
    for _ in range(1500):
        n = int(rng.choice([1, 5, 63, 64, 65, 144, 300]))
        valid, pens = _sequence(rng, n)
        finite = [p for p in pens if np.isfinite(p) and p > 0]
        min_pen0 = F32(rng.choice(finite) if finite and rng.random() < 0.7 else rng.uniform(0.001, 0.3))
        ref = walk_edges(valid, pens, min_pen0)
        assert shared_edges(valid, pens, min_pen0) == ref, (valid, pens, min_pen0)
        kinds.add(ref[0])
    assert kinds == {"sep", "face", "edge"}


def test_edge_face_test_tie_at_rounding():
    # 0.1 and its float32 successor both pass the face test against 0.101 (0.1 * 1.002 + 0.0005 = 0.1007);
    # the first of the smaller wins
    a = F32(0.1)
    b = np.nextafter(a, F32(1))
    pens = [b, a, b, a]
    assert walk_edges([True] * 4, pens, F32(0.101)) == ("edge", 1)
    assert shared_edges([True] * 4, pens, F32(0.101)) == ("edge", 1)
    # a pen whose face test fails keeps the face even when it is the smallest
    assert walk_edges([True], [F32(0.1)], F32(0.1)) == ("face",)
    assert shared_edges([True], [F32(0.1)], F32(0.1)) == ("face",)


def test_shared_face_search_matches_sat_faces():
    rng = np.random.default_rng(11)
    for _ in range(1500):
        n = int(rng.choice([1, 6, 20, 64, 70, 130]))
        _, pens = _sequence(rng, n)
        ref_pen, ref_face = walk_faces(pens)
        pen, face = shared_faces(pens)
        assert face == ref_face
        assert (np.isnan(pen) and np.isnan(ref_pen)) or pen == ref_pen
    # all faces at kDecimalLargest or NaN: face 0 and kDecimalLargest, as the walk leaves them
    assert shared_faces([LARGEST, F32(np.nan)]) == (LARGEST, 0) == walk_faces([LARGEST, F32(np.nan)])
