"""Comparison helpers shared by the parity tests (device step vs reference harness)."""
import numpy as np


def pair_set(c1, c2):
    return {(int(a), int(b)) for a, b in zip(c1, c2)}


def contacts_by_pair(ct):
    out = {}
    for m in range(len(ct["nb_points"])):
        key = (int(ct["colliders"][m, 0]), int(ct["colliders"][m, 1]))
        n = int(ct["nb_points"][m])
        out[key] = dict(n=n, nd=ct["normal_depth"][m, :n], l1=ct["local1"][m, :n], l2=ct["local2"][m, :n],
                        imp=ct["impulse"][m, :n], fric=ct["friction"][m])
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
