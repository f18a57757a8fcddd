"""The wavefront colouring of a one-world group (sv_group_color in csrc/solver.hip) restated on the CPU:
64 manifolds at a time, each starting from the colours earlier chunks left at its bodies and re-choosing
the smallest colour unused by the earlier manifolds of its chunk that share a dynamic body, until no
choice changes.  Its colours must equal sequential greedy colouring in manifold order (the colours of
the one-lane path and the order the solver and the oracle follow), including the clamp to the last
colour when all 64 are used.  Bodies < 0 are static (no constraint)."""
import numpy as np

CHUNK = 64
NCOLORS = 64


def _first_free(used):
    for k in range(NCOLORS):
        if not (used >> k) & 1:
            return k
    return NCOLORS - 1


def greedy(mb, nb):
    mask = [0] * nb
    out = []
    for x, y in mb:
        used = (mask[x] if x >= 0 else 0) | (mask[y] if y >= 0 else 0)
        c = _first_free(used)
        for b in (x, y):
            if b >= 0:
                mask[b] |= 1 << c
        out.append(c)
    return out


def chunked_fixed_point(mb, nb):
    mask = [0] * nb
    out, rounds = [], []
    for i0 in range(0, len(mb), CHUNK):
        ch = mb[i0:i0 + CHUNK]
        pre = [(mask[x] if x >= 0 else 0) | (mask[y] if y >= 0 else 0) for x, y in ch]
        dep = [[k for k in range(j) if (ch[k][0] >= 0 and ch[k][0] in ch[j]) or (ch[k][1] >= 0 and ch[k][1] in ch[j])]
               for j in range(len(ch))]
        c = [_first_free(p) for p in pre]
        for r in range(CHUNK + 1):
            new = []
            for j in range(len(ch)):
                used = pre[j]
                for k in dep[j]:
                    used |= 1 << c[k]
                new.append(_first_free(used))
            changed = new != c
            c = new
            if not changed:
                break
        rounds.append(r + 1)
        for j, (x, y) in enumerate(ch):
            for b in (x, y):
                if b >= 0:
                    mask[b] |= 1 << c[j]
        out += c
    return out, rounds


def _random_manifolds(rng, nb, nm, run_prob):
    mb = []
    for _ in range(nm):
        x, y = int(rng.integers(-1, nb)), int(rng.integers(-1, nb))
        if mb and rng.random() < run_prob:   # runs sharing a body, as the pair order makes them
            x = mb[-1][0]
        if x == y:
            y = -1
        mb.append((x, y))
    return mb


def test_fixed_point_equals_sequential_greedy():
    rng = np.random.default_rng(1)
    worst = 0
    for _ in range(120):
        nb = int(rng.integers(2, 300))
        mb = _random_manifolds(rng, nb, int(rng.integers(1, 400)), 0.5)
        got, rounds = chunked_fixed_point(mb, nb)
        assert got == greedy(mb, nb)
        worst = max(worst, max(rounds))
    assert worst <= CHUNK + 1


def test_dense_conflicts_and_colour_clamp():
    """Few bodies and many manifolds: long in-chunk dependency chains and more than 64 colours wanted
    (the last colour is reused, as the device flags a colour overflow)."""
    rng = np.random.default_rng(2)
    for nb in (2, 3, 5):
        mb = _random_manifolds(rng, nb, 300, 0.9)
        got, _ = chunked_fixed_point(mb, nb)
        assert got == greedy(mb, nb)
    mb = [(0, -1)] * 130   # one body in every manifold: colours 0..63, then the clamp
    got, rounds = chunked_fixed_point(mb, 1)
    assert got == greedy(mb, 1)
    assert got[:NCOLORS] == list(range(NCOLORS)) and all(c == NCOLORS - 1 for c in got[NCOLORS:])
    assert max(rounds) <= CHUNK + 1
