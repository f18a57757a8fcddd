"""The two modes the throughput figure is measured in, checked against the coloured-order oracle on the
full batch of 4096 uneven worlds (8-512 bodies, about 500k bodies): every world iterates its velocity
solver to the residual 1e-4 (cap 100), with the plain warm start and with the learned one.

Each step, the oracle (one reference PhysicsWorld per world, stepped on host threads) is fed each
world's manifold order, and with the learned warm start the network's predicted impulses of the
device's manifolds; it applies its own stopping rule.  Checked every step: each world's sweep count
equals the oracle's, oriented pair sets and per-manifold point counts are equal, contacts agree within
1e-5 relative and body states within 1e-4.  The network here is untrained (seeded weights, modest output
scale): the check is how predictions are applied, not their quality, which tests/test_gpu_mlp.py
measures."""
import time

import numpy as np
import pytest

from reactphysics3d_amd import scene as S
from tests.parity import compare_contacts_batch, same_pair_sets

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
RESIDUAL, CAP = 1e-4, 100
STEPS = 32   # stacks collide from the first step, piles land after about 25 steps


@pytest.fixture(scope="module")
def worlds():
    return S.batched_worlds(4096, seed=4)


@pytest.mark.parametrize("warm", ["plain", "learned"])
def test_full_batch_residual_mode_matches_oracle(worlds, warm):
    from oracle.ref import RefBatch
    from reactphysics3d_amd.lib import Batch, Mlp
    t0 = time.time()
    dev = Batch(worlds, solver="fused")
    ref = RefBatch(worlds)
    order = dev.joint_order()
    ref.set_joint_orders([order[(order >= dev.joint_base[w]) & (order < dev.joint_base[w + 1])] - dev.joint_base[w]
                          for w in range(len(worlds))])
    dev.set_residual_mode(RESIDUAL, CAP)
    mlp = None
    if warm == "learned":
        mlp = Mlp(seed=5)
        mlp.set_normalization(np.zeros(24, np.float32), np.full(24, 0.5, np.float32),
                              np.array([0.2, 0.05, 0.05], np.float32))
        dev.set_warm_start("learned", mlp)
    cbase = np.asarray(dev.collider_base, np.int64)
    print(f"[{warm}] set-up {time.time() - t0:.1f}s, {dev.nb} bodies", flush=True)
    worst, predicted, sweeps = 0.0, 0, []
    for k in range(STEPS):
        dev.step(DT)
        ct = dev.contacts()
        wm = np.searchsorted(cbase, ct["colliders"][:, 0], side="right") - 1
        pred = None
        if warm == "learned":
            pred = dev.warm_start_predictions().reshape(-1, 12)
            predicted += int(np.count_nonzero(pred[:, 0::3]))
        it_ref, res_ref = ref.step_ordered(DT, ct["colliders"][:, 0], ct["colliders"][:, 1], wm, RESIDUAL, CAP, pred)
        it_dev, res_dev = dev.world_iterations()
        bad = np.nonzero(it_dev != it_ref)[0]
        assert len(bad) == 0, (f"step {k}: {len(bad)} worlds ran a different sweep count, e.g. world {bad[0]}: "
                               f"{it_dev[bad[0]]} vs oracle {it_ref[bad[0]]}")
        assert np.array_equal(res_dev, res_ref), f"step {k}: final residuals differ"
        p1, p2 = dev.pairs()
        r1, r2 = ref.pairs()
        assert same_pair_sets(p1, p2, r1, r2), f"step {k}: pair sets differ"
        errs = compare_contacts_batch(ct, ref.contacts())
        assert not errs, f"step {k}: " + "; ".join(errs[:5])
        d = np.concatenate([a.reshape(len(a), -1) for a in dev.bodies()], axis=1)
        r = np.concatenate([a.reshape(len(a), -1) for a in ref.bodies()], axis=1)
        err = float(np.max(np.abs(d - r)))
        worst = max(worst, err)
        assert err <= 1e-4, f"step {k}: body state differs by {err}"
        sweeps.append(float(it_dev.mean()))
    print(f"[{warm}] {STEPS} steps in {time.time() - t0:.1f}s: worst body diff {worst}, mean sweeps per step "
          f"{np.round(sweeps, 1).tolist()}, predicted points {predicted}", flush=True)
    assert max(sweeps) > 10, "the residual rule never ran past the fixed iteration count"
    if warm == "learned":
        assert predicted > 0, "no prediction reached the solver"
    ref.close()
    dev.close()
    if mlp is not None:
        mlp.close()
