"""FP32 error of the device step against the FP64 build of the reference, stated per workload.

The device reproduces the FP32 reference bit for bit (the parity tests), so its error against FP64 is
the FP32 reference's own rounding, fed the same solve order.  Each workload is stepped on the device and
in the FP64 reference (device contact and joint order) until the FP64 world stops following the FP32 one
(a manifold of the device's order is missing from the FP64 world, or the pair sets part): piles and
falling containers are chaotic, so the horizon is finite and the error grows along it.  The test prints
one "FP64_ERROR {json}" line per workload (per-step maxima of |dposition|, |dorientation| and
|dvelocity|) and checks that the first steps agree to FP32 rounding."""
import json

import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import same_pair_sets

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _f64_available():
    from oracle import ref
    return ref.available(f64=True)


WORKLOADS = {
    "box_pile_1000": (lambda: [S.box_pile(10, 10, seed=1)], 120),
    "mixed_container_65536": (lambda: [S.mixed_container(65536, seed=2)], 20),
    "hulls_16384_chains_256x16": (lambda: [S.hull_scene(16384, 256, 16, seed=3)], 20),
    "batched_256_of_4096_worlds": (lambda: S.batched_worlds(4096, seed=4)[::16], 60),
}


def _state_errors(dev, dstate, w, ref):
    b0, b1 = dev.body_base[w], dev.body_base[w + 1]
    pd, qd, vd, wd = (a[b0:b1] for a in dstate)
    pr, qr, vr, wr = ref.bodies()
    return (float(np.abs(pd - pr).max()), float(np.abs(qd - qr).max()),
            float(max(np.abs(vd - vr).max(), np.abs(wd - wr).max())))


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_fp32_error_against_fp64_reference(name):
    if not _f64_available():
        pytest.skip("FP64 reference not built (make -C oracle f64)")
    from oracle.ref import RefWorld
    from reactphysics3d_amd.lib import Batch
    make, steps = WORKLOADS[name]
    scenes = make()
    dev = Batch(scenes)
    refs = [RefWorld(sc, f64=True) for sc in scenes]
    order = dev.joint_order()
    for w, ref in enumerate(refs):
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        ref.set_joint_order(order[(order >= j0) & (order < j1)] - j0)
    alive = np.ones(len(scenes), bool)
    horizon = np.zeros(len(scenes), np.int64)
    per_step = []   # (max |dpos|, max |dq|, max |dvel|) over the worlds still followed
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        p1, p2 = dev.pairs()
        dstate = dev.bodies()
        worst = [0.0, 0.0, 0.0]
        for w, ref in enumerate(refs):
            if not alive[w]:
                continue
            c0, c1 = dev.collider_base[w], dev.collider_base[w + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            try:
                ref.step_ordered(DT, ct["colliders"][sel, 0] - c0, ct["colliders"][sel, 1] - c0)
            except RuntimeError:
                alive[w] = False
                continue
            ps = (p1 >= c0) & (p1 < c1)
            r1, r2 = ref.pairs()
            if not same_pair_sets(p1[ps] - c0, p2[ps] - c0, r1, r2):
                alive[w] = False
                continue
            horizon[w] = k + 1
            e = _state_errors(dev, dstate, w, ref)
            worst = [max(a, b) for a, b in zip(worst, e)]
        if not alive.any():
            break
        per_step.append(worst)
    err = np.array(per_step) if per_step else np.zeros((0, 3))
    out = {"workload": name, "worlds": len(scenes), "steps_run": steps,
           "median_horizon_steps": float(np.median(horizon)), "min_horizon_steps": int(horizon.min()),
           "first_step": [float(x) for x in err[0]] if len(err) else None,
           "step_10": [float(x) for x in err[9]] if len(err) >= 10 else None,
           "max_over_horizon": [float(x) for x in err.max(0)] if len(err) else None,
           "columns": ["max |dposition| m", "max |dorientation|", "max |dvelocity| m/s or rad/s"]}
    print("FP64_ERROR " + json.dumps(out), flush=True)
    assert len(err) >= 3 and horizon.max() >= 3
    # one step from identical states differs by FP32 rounding only
    assert err[0][0] <= 1e-4 and err[0][1] <= 1e-4 and err[0][2] <= 1e-3
