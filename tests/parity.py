"""Comparison helpers shared by the parity tests (device step vs reference harness)."""
import numpy as np


def pair_set(c1, c2):
    return {(int(a), int(b)) for a, b in zip(c1, c2)}


def contacts_by_pair(ct):
    """Manifolds keyed by (collider 1, collider 2), and by their index among the pair's manifolds when the
    pair has several (a convex vs concave pair has up to four)."""
    out = {}
    sub = ct.get("manifold")
    for m in range(len(ct["nb_points"])):
        key = (int(ct["colliders"][m, 0]), int(ct["colliders"][m, 1]))
        if sub is not None and int(sub[m]) > 0:
            key = key + (int(sub[m]),)
        n = int(ct["nb_points"][m])
        out[key] = dict(n=n, nd=ct["normal_depth"][m, :n], l1=ct["local1"][m, :n], l2=ct["local2"][m, :n])
    return out


def max_rel(a, b, floor=1.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def compare_contacts(dev, ref, tol=1e-5):
    """Returns a list of mismatch descriptions (empty when the manifolds agree)."""
    d = contacts_by_pair(dev)
    r = contacts_by_pair(ref)
    errs = []
    if set(d) != set(r):
        errs.append(f"manifold pair sets differ: dev-only {sorted(set(d) - set(r))[:5]} "
                    f"ref-only {sorted(set(r) - set(d))[:5]}")
    for k in set(d) & set(r):
        if d[k]["n"] != r[k]["n"]:
            errs.append(f"pair {k}: {d[k]['n']} points vs reference {r[k]['n']}")
            continue
        for f in ("nd", "l1", "l2"):
            e = max_rel(d[k][f], r[k][f])
            if e > tol:
                errs.append(f"pair {k}: {f} differs by {e:.3e}")
    return errs


def _pair_keys(c1, c2):
    return (np.asarray(c1, np.int64) << 32) | np.asarray(c2, np.int64)


def same_pair_sets(c1, c2, r1, r2):
    """Oriented pair sets equal (vectorised pair_set comparison)."""
    a = np.sort(_pair_keys(c1, c2))
    b = np.sort(_pair_keys(r1, r2))
    return a.shape == b.shape and bool(np.array_equal(a, b))


def compare_contacts_batch(dev, ref, tol=1e-5):
    """compare_contacts over many manifolds at once (batch numbering on both sides)."""
    kd = _pair_keys(dev["colliders"][:, 0], dev["colliders"][:, 1])
    kr = _pair_keys(ref["colliders"][:, 0], ref["colliders"][:, 1])
    od, orf = np.argsort(kd, kind="stable"), np.argsort(kr, kind="stable")
    if kd.shape != kr.shape or not np.array_equal(kd[od], kr[orf]):
        sd, sr = set(kd.tolist()), set(kr.tolist())
        return [f"manifold pair sets differ: {len(sd - sr)} device-only, {len(sr - sd)} reference-only "
                f"(e.g. {[(k >> 32, k & 0xFFFFFFFF) for k in sorted(sd ^ sr)[:4]]})"]
    nd_, nr_ = dev["nb_points"][od], ref["nb_points"][orf]
    errs = []
    bad = np.nonzero(nd_ != nr_)[0]
    if len(bad):
        errs.append(f"{len(bad)} manifolds with different point counts, e.g. pair "
                    f"{tuple(dev['colliders'][od[bad[0]]])}: {nd_[bad[0]]} vs {nr_[bad[0]]}")
        return errs
    live = np.arange(4)[None, :] < nd_[:, None]
    for f in ("normal_depth", "local1", "local2"):
        a = np.asarray(dev[f][od], np.float64)
        b = np.asarray(ref[f][orf], np.float64)
        e = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
        e = np.where(live[:, :, None], e, 0.0)
        if e.size and float(e.max()) > tol:
            i = int(np.argmax(e.max(axis=(1, 2))))
            errs.append(f"{f} differs by {float(e.max()):.3e} (pair {tuple(dev['colliders'][od[i]])})")
    return errs
