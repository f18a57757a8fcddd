#!/usr/bin/env python3
"""Benchmark of the MI355X PhysicsWorld::update() step path.

Default workload (one MI355X): 4096 independent worlds in one batch, body counts
log-uniform in [8, 512] (about 500k bodies), each a seeded box pile or box stack
(height 4-20) with 0-32 hinge / fixed joints on a static ground, dt = 1/60 s,
FP32, stepped by a captured hipGraph, every world iterating its velocity solver
to the residual 1e-4 (iteration cap 100).  Two copies of the batch are timed from
the same warmed-up state: plain warm start (cached impulses, zero for new
contacts, as the reference) and learned warm start (an MLP trained just before,
on the GPU, on converged impulses logged from the plain solver on 1000 held-out
worlds).  `value` is the learned form; both are reported.  Inputs are resident
in HBM before the timed regions.  A step is one full PhysicsWorld::update() of
every world: broadphase, narrowphase, contact creation, coloured contact + joint
solver, integration, refit.

Other workloads: --workload mixed (65,536 spheres/boxes/capsules in a walled
container, 6/3 iterations), --workload pile (1000 boxes, 10/5 iterations) and
--workload hulls (16,384 convex hulls of 24 ellipsoid points + 256 chains of 16
spheres on ball-and-socket joints, 6/3 iterations).

Prints one JSON line (see the repository's bench contract); the per-kernel
breakdown goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, vendor figure
DT = 1.0 / 60.0
RESIDUAL = 1e-4
RESIDUAL_CAP = 100


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic (compulsory) HBM bytes per launch: every datum a launch must read or write, once.
# c = counts of a step, nj = joints, ndyn = dynamic bodies.
def bytes_per_launch(name, c, ndyn, nj):
    nt = c["nb_tests"]
    m, p = c["nb_manifolds"], c["nb_contact_points"]
    ncol = max(1, c["nb_colors"])
    nb, ncl, npairs = c["nb_bodies"], c["nb_colliders"], c["nb_pairs"]
    np_types = {"np_sphere_sphere": 1, "np_sphere_capsule": 2, "np_capsule_capsule": 3, "np_sphere_poly": 4,
                "np_capsule_poly": 5, "np_poly_poly": 6}
    if name == "np_box_box":
        tests = c["nb_box_tests"]
        return 176 * tests + 40 * p * (tests / max(1, sum(nt)))
    if name == "np_poly_poly":
        # pairs with a convex mesh: bytes of the pair records (the hull data is L2 resident); the kernel is
        # rooflined with the VALU FLOP model instead (poly_flops_per_launch)
        tests = nt[6] - c["nb_box_tests"]
        return 176 * tests + 40 * p * (tests / max(1, sum(nt)))
    if name in np_types:
        tests = nt[np_types[name]]
        return 176 * tests + 40 * p * (tests / max(1, sum(nt)))
    if name == "sv_world_solve":
        # whole velocity solver of a step, one launch: 48 B / contact point + 88 B / manifold +
        # 140 B / dynamic body + 200 B / joint
        return 48 * p + 88 * m + 140 * ndyn + 200 * nj
    if name == "world_position":
        # position integration + commit (192 B / dynamic body) and fat-AABB refit (48 B / collider)
        return 192 * ndyn + 48 * ncl
    if name in ("sv_solve_color", "sv_warmstart"):
        return (48 * p + 88 * m + 140 * 2 * m) / ncol
    if name == "sv_prepare":
        return 420 * m + 160 * p
    if name == "bp_traverse":
        return 48 * ncl + 4 * ncl + 8 * npairs / 2
    if name == "bp_merge":
        return 112 * npairs   # keys, pair records and the link to the previous cache; caches are not copied
    if name == "dyn_integrate_velocities":
        return nb * (16 * 9 + 16 * 5)
    if name == "dyn_integrate_positions":
        return nb * (16 * 6 + 16 * 2)
    if name == "dyn_update_state":
        return nb * (16 * 5 + 16 * 7)
    if name == "bp_refit":
        return ncl * (16 * 6 + 16 * 6 + 8)
    return None


def hull_stats(scenes):
    """Mean faces, vertices and edges of the convex meshes of the workload (None without meshes)."""
    f = v = e = n = 0
    for sc in scenes:
        for verts, faces in getattr(sc, "hulls", []):
            f += len(faces)
            v += len(verts)
            e += sum(len(x) for x in faces) // 2
            n += 1
    return (f / n, v / n, e / n) if n else None


def poly_flops_per_launch(c, hs):
    """Polyhedron SAT FLOP model per test, 6 (F1 V2 + F2 V1) + 30 E1 E2 (face support searches and the
    edge-pair search), with the workload's mean mesh for both shapes, times the mesh tests of a step."""
    F, V, E = hs
    tests = c["nb_tests"][6] - c["nb_box_tests"]
    return tests * (6.0 * (F * V + F * V) + 30.0 * E * E)


MLP_INFER_FLOP = 39680      # 2 (24 x 128 + 128 x 128 + 128 x 3) per inferred contact
MLP_TRAIN_FLOP = 3 * 39680  # forward + backward per training sample
MFMA_BF16_PEAK = 2.5e15     # dense BF16 MFMA, MI355X vendor figure
VALU_FP32_PEAK = 157.3e12   # FP32 vector, MI355X vendor figure

BYTE_MODEL = {
    "sv_world_solve": "48 B/contact point + 88 B/manifold + 140 B/dynamic body + 200 B/joint per launch (one step)",
    "world_position": "192 B/dynamic body + 48 B/collider per launch (one step)",
    "np_poly_poly": "176 B/test + 40 B/contact point",
    "np_box_box": "176 B/test + 40 B/contact point",
}

KERNELS = ["bp_bounds", "bp_morton", "bp_sort", "bp_karras", "bp_fit", "bp_world_roots", "bp_traverse", "bp_copy_pairs",
           "bp_walk_again", "bp_world_scan",
           "bp_merge",
           "np_classify", "np_sphere_sphere", "np_sphere_capsule", "np_capsule_capsule", "np_sphere_poly",
           "np_capsule_poly", "np_box_box", "np_poly_poly", "sv_flag", "sv_compact", "sv_adj_count", "sv_adj_fill",
           "sv_color_round", "sv_color_finish", "sv_color_keys", "sv_color_sort", "sv_world_ranges", "sv_group_color",
           "sv_world_prepare", "mlp_infer", "sv_world_solve", "sv_log_samples", "world_position",
           "dyn_integrate_velocities", "sv_prepare", "sv_warmstart", "sv_solve_color", "sv_store", "jt_init",
           "jt_warmstart", "jt_solve_velocity", "jt_solve_position", "dyn_integrate_positions", "dyn_update_state",
           "bp_refit"]


def build_workload(args):
    from reactphysics3d_amd import scene as S
    if args.workload == "batched":
        scenes = S.batched_worlds(args.worlds, seed=4)
        name = f"batched_{args.worlds}_worlds_8_512_bodies"
    elif args.workload == "mixed":
        scenes = [S.mixed_container(args.bodies, seed=2)]
        name = f"mixed_container_{args.bodies}"
    elif args.workload == "pile":
        scenes = [S.box_pile(10, 10, seed=1)]
        name = "box_pile_1000"
    elif args.workload == "hulls":
        scenes = [S.hull_scene(16384, 256, 16, seed=3)]
        name = "hulls_16384_chains_256x16"
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    return name, scenes


def breakdown(dev, steps, tag):
    out = {}
    for k in KERNELS:
        ms, n = dev.kernel_time(k)
        if n:
            out[k] = (ms, n)
    total = sum(ms * n for ms, n in out.values())
    for k, (ms, n) in sorted(out.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        log(f"[bench] {tag} {k:26s} {ms * 1e3:9.1f} us x {n / steps:6.1f}/step  {100 * ms * n / max(total, 1e-12):5.1f}%")
    return out, total


def cpu_baseline(scenes, args, ndyn_total):
    """Reference PhysicsWorld::update() on the host cores: a bounded sample of the worlds, one world per
    thread at a time (worlds are created serially), fixed WorldSettings iteration counts."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ref
    if not ref.available():
        log("[bench] reference harness not built; no CPU baseline")
        return None
    # one world per host thread; 16 threads = the host CPU share of one GPU on the GPU box (the box
    # reports more CPUs than a single-GPU job may use)
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    stride = max(1, len(scenes) // args.cpu_worlds) if len(scenes) > 1 else 1
    sample = scenes[::stride][: args.cpu_worlds] if len(scenes) > 1 else scenes
    if len(sample) == 1 and len(scenes) == 1:
        secs, cts = ref.cpu_bench(sample[0], threads, 0, args.cpu_steps, DT)
        nd = sample[0].nb_dynamic
        return {"value": round(nd * args.cpu_steps * threads / secs, 1), "unit": "body-steps/s", "cores": threads,
                "host_cpus": os.cpu_count(), "kind": "reference", "contacts_per_s": round(cts / secs, 1),
                "sample": f"{threads} threads x 1 copy each of the {sample[0].nb_bodies}-body world, first "
                          f"{args.cpu_steps} steps, ReactPhysics3D 0.10.2 g++ -O3 FP32, fixed iterations"}
    worlds = [ref.RefWorld(sc) for sc in sample]
    nd = sum(sc.nb_dynamic for sc in sample)

    def run(w, n):
        for _ in range(n):
            w.step(DT)

    with ThreadPoolExecutor(max_workers=threads) as ex:   # same untimed warmup as the device
        list(ex.map(lambda w: run(w, args.warmup), worlds))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda w: run(w, args.cpu_steps), worlds))
    secs = time.perf_counter() - t0
    for w in worlds:
        w.close()
    return {"value": round(nd * args.cpu_steps / secs, 1), "unit": "body-steps/s", "cores": threads,
            "host_cpus": os.cpu_count(), "kind": "reference",
            "sample": f"{len(sample)} of the {len(scenes)} worlds (every {stride}th, {nd} dynamic bodies), "
                      f"{args.cpu_steps} steps each after {args.warmup} untimed, one world per thread at a time on "
                      f"{threads} threads, "
                      f"ReactPhysics3D 0.10.2 g++ -O3 FP32, fixed 10/5 iterations"}


def train_warm_start(args):
    """Learned warm start: (features, converged impulses) logged from the plain solver at the residual on
    held-out worlds (stream seeds 1000-1999), network trained on the GPU by the project's kernels."""
    from reactphysics3d_amd import scene as S
    from reactphysics3d_amd.lib import Batch, Mlp
    t0 = time.time()
    held = Batch(S.batched_worlds(args.train_worlds, seed=1000))
    held.set_residual_mode(RESIDUAL, RESIDUAL_CAP)
    held.set_logging(args.train_samples)
    steps = 0
    while steps < args.train_max_steps:
        held.step(DT, 10)
        steps += 10
        if log_count(held) >= args.train_samples:
            break
    F, T = held.get_log(args.train_samples)
    held.close()
    t1 = time.time()
    mlp = Mlp(seed=7)
    mlp.set_normalization(F.mean(0), 1.0 / (F.std(0) + 1e-6), T.std(0) + 1e-6)
    loss = mlp.train(F, T, epochs=args.train_epochs, minibatch=4096, lr=2e-3, seed=11)
    t2 = time.time()
    flop = float(len(F)) * args.train_epochs * MLP_TRAIN_FLOP
    info = {"samples": int(len(F)), "logged_steps": steps, "held_out_worlds": args.train_worlds,
            "train_tflop_s": round(flop / max(1e-9, t2 - t1) / 1e12, 3),
            "train_mfma_frac_whole_call": round(flop / max(1e-9, t2 - t1) / MFMA_BF16_PEAK, 6),
            "epochs": args.train_epochs, "minibatch": 4096, "final_loss": round(float(loss), 5),
            "log_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}
    log(f"[bench] warm-start network: {info}")
    return mlp, info


def log_count(batch):
    from reactphysics3d_amd import lib
    return int(lib.load().rp3dg_batch_get_log(batch.h, None, None, 1 << 62))


def fp64_leg(scenes, steps=20, sample=64):
    """FP32 device error against the FP64 reference fed the device's order, on a sample of the worlds:
    per-step maxima of |dposition| and |dvelocity| while the FP64 world follows (its manifolds are the
    device's).  None when the FP64 reference is not built."""
    from oracle import ref
    if not ref.available(f64=True):
        return None
    from reactphysics3d_amd.lib import Batch
    stride = max(1, len(scenes) // sample)
    sub = scenes[::stride][:sample]
    dev = Batch(sub)
    refs = [ref.RefWorld(sc, f64=True) for sc in sub]
    order = dev.joint_order()
    for w, r in enumerate(refs):
        j0, j1 = dev.joint_base[w], dev.joint_base[w + 1]
        r.set_joint_order(order[(order >= j0) & (order < j1)] - j0)
    import numpy as np
    alive = np.ones(len(sub), bool)
    horizon = np.zeros(len(sub), int)
    curve = []
    for k in range(steps):
        dev.step(DT)
        ct = dev.contacts()
        st = dev.bodies()
        worst = [0.0, 0.0]
        for w, r in enumerate(refs):
            if not alive[w]:
                continue
            c0, c1 = dev.collider_base[w], dev.collider_base[w + 1]
            sel = (ct["colliders"][:, 0] >= c0) & (ct["colliders"][:, 0] < c1)
            try:
                r.step_ordered(DT, ct["colliders"][sel, 0] - c0, ct["colliders"][sel, 1] - c0)
            except RuntimeError:
                alive[w] = False
                continue
            horizon[w] = k + 1
            b0, b1 = dev.body_base[w], dev.body_base[w + 1]
            rp, rq, rv, rw = r.bodies()
            worst[0] = max(worst[0], float(np.abs(st[0][b0:b1] - rp).max()), float(np.abs(st[1][b0:b1] - rq).max()))
            worst[1] = max(worst[1], float(np.abs(st[2][b0:b1] - rv).max()), float(np.abs(st[3][b0:b1] - rw).max()))
        curve.append(worst)
        if not alive.any():
            break
    for r in refs:
        r.close()
    dev.close()
    return {"sample": f"{len(sub)} of {len(scenes)} worlds (every {stride}th), {steps} steps, fixed iterations, "
                      f"FP64 reference fed the device's contact and joint order",
            "first_step": [float(f"{x:.3g}") for x in curve[0]] if curve else None,
            "step_10": [float(f"{x:.3g}") for x in curve[9]] if len(curve) >= 10 else None,
            "max": [float(f"{max(c[i] for c in curve):.3g}") for i in range(2)] if curve else None,
            "median_horizon_steps": float(np.median(horizon)),
            "columns": ["max |d position or orientation|", "max |d linear or angular velocity|"]}


def iteration_curves(scenes, mlp, steps):
    """Plain and learned warm start stepped from the same initial state (fresh batches, residual mode,
    eager): mean velocity sweeps and new contact points per step, so the early steps where contacts are
    new show next to the settled ones."""
    from reactphysics3d_amd.lib import Batch
    out = {}
    for tag in ("plain", "learned"):
        d = Batch(scenes)
        d.set_residual_mode(RESIDUAL, RESIDUAL_CAP)
        if tag == "learned":
            d.set_warm_start("learned", mlp)
        else:
            d.set_logging(1)   # counts the new points per step (one sample kept)
        its, new = [], []
        for _ in range(steps):
            d.step(DT)
            its.append(d.iteration_stats()["mean_iterations"])
            new.append(d.counts()["nb_new_points"])
        out[tag] = (its, new)
        d.close()
    p_its, p_new = out["plain"]
    l_its, _ = out["learned"]
    ratio = [l / max(1e-9, p) for l, p in zip(l_its, p_its)]
    k = min(30, steps)
    return {"steps": steps, "plain_mean_sweeps": [round(x, 2) for x in p_its],
            "learned_mean_sweeps": [round(x, 2) for x in l_its], "new_points": [int(x) for x in p_new],
            "ratio_per_step": [round(x, 4) for x in ratio],
            "ratio_first_30": round(sum(l_its[:k]) / max(1e-9, sum(p_its[:k])), 4),
            "ratio_all": round(sum(l_its) / max(1e-9, sum(p_its)), 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="batched", choices=["batched", "mixed", "pile", "hulls"])
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--bodies", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP32-vs-FP64 error leg")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-worlds", type=int, default=512)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-profile", action="store_true", help="no HIP event timing at all")
    ap.add_argument("--no-tune", action="store_true", help="skip the solver launch-shape autotuner")
    ap.add_argument("--group-bodies", type=int, default=0,
                    help="fused solver: pack consecutive worlds into workgroups of up to this many bodies "
                         "(0 = library default, 1 = one world per workgroup)")
    ap.add_argument("--train-worlds", type=int, default=1000)
    ap.add_argument("--train-samples", type=int, default=2_000_000)
    ap.add_argument("--train-max-steps", type=int, default=400)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--curve-steps", type=int, default=60,
                    help="steps of the plain / learned iteration curves from the initial state (0 = off)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run a plain elementwise copy of MIB MiB first (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {args.copy_probe} MiB elementwise copies")
        del src, dst

    from reactphysics3d_amd.lib import Batch

    t0 = time.time()
    name, scenes = build_workload(args)
    ndyn = sum(sc.nb_dynamic for sc in scenes)
    nj = sum(len(sc.joints) for sc in scenes)
    hs = hull_stats(scenes)
    log(f"[bench] workload {name}: {len(scenes)} worlds, {sum(sc.nb_bodies for sc in scenes)} bodies, "
        f"{ndyn} dynamic, {nj} joints, built in {time.time() - t0:.1f}s")
    t0 = time.time()
    batched = args.workload == "batched"
    dev = Batch(scenes, group_bodies=args.group_bodies)
    dev_l = Batch(scenes, group_bodies=args.group_bodies) if batched else None   # same worlds, learned warm start
    log(f"[bench] uploaded in {time.time() - t0:.1f}s")
    for d in (dev, dev_l):
        if d is None:
            continue
        for _ in range(args.warmup):
            d.step(DT)
        d.sync()
    log(f"[bench] warmup {args.warmup} steps done; counts {dev.counts()}")
    graph = batched or len(scenes) > 1
    if not graph:
        # a single world in the fused layout replays captured steps as well (the per-colour layout of
        # large worlds reads its colour count back every step and runs eagerly)
        try:
            dev.step_graph(DT, 1)
            dev.sync()
            graph = True
        except RuntimeError:
            pass
    training = None
    mlp = None
    if batched:
        mlp, training = train_warm_start(args)
        for d in (dev, dev_l):
            d.set_residual_mode(RESIDUAL, RESIDUAL_CAP)
        dev_l.set_warm_start("learned", mlp)

    tuned = None
    if not args.no_tune:
        # autotuner: solver launch shape timed over real steps in the timed configuration (residual mode,
        # plain warm start); the learned copy runs with the same shape so that the two differ only in the
        # warm start
        choice, ms = dev.autotune(DT, 2)
        tuned = {"choice": choice, "ms_per_step": round(ms, 3), "trials": [[c, round(t, 3)] for c, t in dev.tune_table()],
                 "meaning": ("world-class workgroup = clamp(bodies bound / choice, 64, 256)" if graph
                             else "threads per block of the colour kernels"),
                 "applied_to": "plain and learned"}
        if dev_l is not None:
            dev_l.set_solver_shape(choice)
        log(f"[bench] autotuned solver launch shape: {tuned}")

    def timed(d, tag):
        c0 = d.counts()
        d.sync()
        t0 = time.perf_counter()
        if graph:
            d.step_graph(DT, args.steps)
        else:
            d.step(DT, args.steps)
        d.sync()
        wall = time.perf_counter() - t0
        c1 = d.counts()
        r = {"body_steps_per_s": round(ndyn * args.steps / wall, 1), "ms_per_step": round(wall * 1e3 / args.steps, 4),
             "contacts_per_s": round((c1["total_contact_points"] - c0["total_contact_points"]) / wall, 1),
             "wall": wall, "counts": c1}
        if batched:
            its, res = d.world_iterations()
            r["last_step"] = {"mean_velocity_iterations": round(float(its.mean()), 2),
                              "worlds_at_iteration_cap": int((its >= RESIDUAL_CAP).sum()),
                              "max_final_residual": float(f"{float(res.max()):.3g}"),
                              "max_final_residual_below_cap": float(f"{float(res[its < RESIDUAL_CAP].max()):.3g}")
                              if (its < RESIDUAL_CAP).any() else None,
                              "new_point_fraction": round(c1["nb_new_points"] / max(1, c1["nb_contact_points"]), 4)}
        log(f"[bench] timed {tag} {args.steps} steps ({'graph' if graph else 'eager'}): {r['ms_per_step']:.3f} ms/step, "
            f"{r['body_steps_per_s']:.4g} body-steps/s, {r['contacts_per_s']:.4g} contacts/s, {r.get('last_step')}")
        return r

    # ---------------- timed regions
    plain = timed(dev, "plain")
    learned = timed(dev_l, "learned") if batched else None
    head = learned if batched else plain
    main_dev = dev_l if batched else dev
    wall, c1 = head["wall"], head["counts"]
    value = head["body_steps_per_s"]

    # mean sweeps per step over an eager pass of the same length (per-step stats, both copies), untimed
    sweeps = None
    if batched:
        sweeps = {}
        for tag, d in (("plain", dev), ("learned", dev_l)):
            acc = []
            for _ in range(min(args.steps, 50)):
                d.step(DT)
                acc.append(d.iteration_stats()["mean_iterations"])
            sweeps[tag] = round(sum(acc) / len(acc), 2)
        plain["mean_velocity_iterations"] = sweeps["plain"]
        learned["mean_velocity_iterations"] = sweeps["learned"]

    # per-kernel breakdown (eager, every kernel timed) and the dominant kernel
    roofline, kernels = None, None
    if not args.no_profile:
        main_dev.profile(True)
        nprof = 5
        main_dev.step(DT, nprof)
        items, _ = breakdown(main_dev, nprof, "eager")
        cp = main_dev.counts()
        main_dev.profile(False)
        kernels = {}
        for k, (ms, n) in items.items():
            b = bytes_per_launch(k, cp, ndyn, nj)
            e = {"mean_us": round(ms * 1e3, 2), "launches_per_step": round(n / nprof, 2)}
            if b is not None:
                e["gb_s"] = round(b / (ms * 1e-3) / 1e9, 2)
                e["hbm_frac"] = round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
            if k == "np_poly_poly" and hs:
                fl = poly_flops_per_launch(cp, hs)
                e["tflop_s"] = round(fl / (ms * 1e-3) / 1e12, 3)
                e["valu_frac"] = round(fl / (ms * 1e-3) / VALU_FP32_PEAK, 5)
            if k == "mlp_infer":
                fl = MLP_INFER_FLOP * cp["nb_new_points"]
                e["tflop_s"] = round(fl / (ms * 1e-3) / 1e12, 3)
                e["mfma_frac"] = round(fl / (ms * 1e-3) / MFMA_BF16_PEAK, 6)
            kernels[k] = e
        dominant = None
        for k, _ in sorted(items.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
            if bytes_per_launch(k, cp, ndyn, nj) is not None:
                dominant = k
                break
        if dominant:
            # graph replays are not event-timed: the dominant kernel is timed over an eager pass of the same
            # number of steps right after the timed regions (two HIP events per step, batch stream)
            main_dev.profile(True, only=dominant)
            main_dev.step(DT, args.steps)
            main_dev.sync()
            ms, n = main_dev.kernel_time(dominant)
            main_dev.profile(False)
            cd = main_dev.counts()
            traffic, traffic_src = None, None
            tpath = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
            if os.path.exists(tpath):
                with open(tpath) as f:
                    tj = json.load(f)
                traffic = tj.get(dominant)
                traffic_src = tj.get("_measured")
            if dominant == "np_poly_poly" and hs:
                fl = poly_flops_per_launch(cd, hs)
                achieved = fl / (ms * 1e-3) / 1e12
                roofline = {"kernel": dominant, "bound": "valu", "achieved": round(achieved, 3),
                            "peak": VALU_FP32_PEAK / 1e12, "unit": "TFLOP/s",
                            "frac": round(achieved * 1e12 / VALU_FP32_PEAK, 5), "traffic": traffic,
                            "flop_model": "6 (F1 V2 + F2 V1) + 30 E1 E2 per mesh test, mean mesh "
                                          f"F {hs[0]:.1f} V {hs[1]:.1f} E {hs[2]:.1f}"}
            else:
                b = bytes_per_launch(dominant, cd, ndyn, nj)
                achieved = b / (ms * 1e-3) / 1e9
                roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                            "bytes_per_launch": int(b), "byte_model": BYTE_MODEL.get(dominant, "")}
            roofline.update({"traffic_source": traffic_src, "mean_launch_us": round(ms * 1e3, 2), "launches": n,
                             "timing": "HIP events around the kernel on the batch's stream, eager pass of the same "
                                       "number of steps directly after the timed regions",
                             "share_of_step": round(ms / (wall * 1e3 / args.steps), 3)})
            if dominant == "sv_world_solve" and learned and learned.get("mean_velocity_iterations"):
                # every sweep re-reads the rows: 80 B / point (normal, r1, r2, impulse r/w) + 124 B / manifold
                sweep = 80 * cd["nb_contact_points"] + 124 * cd["nb_manifolds"]
                its = learned["mean_velocity_iterations"]
                roofline["row_bytes_per_sweep"] = int(sweep)
                roofline["row_traffic_gbs"] = round(sweep * its / (ms * 1e-3) / 1e9, 2)

    curves = iteration_curves(scenes, mlp, args.curve_steps) if batched and args.curve_steps else None

    fixed = None
    if batched:
        # the same worlds with the fixed WorldSettings iteration count (the reference's own mode)
        dev.set_residual_mode(0.0)
        k = max(10, args.steps // 4)
        dev.sync()
        t0 = time.perf_counter()
        dev.step_graph(DT, k)
        dev.sync()
        w2 = time.perf_counter() - t0
        fixed = {"body_steps_per_s": round(ndyn * k / w2, 1), "ms_per_step": round(w2 * 1e3 / k, 4),
                 "velocity_iterations": int(scenes[0].settings["vel_iters"]), "steps": k}
        log(f"[bench] fixed iterations: {w2 * 1e3 / k:.3f} ms/step, {ndyn * k / w2:.4g} body-steps/s")

    fp64 = None
    if not args.no_fp64:
        try:
            fp64 = fp64_leg(scenes, steps=20)
            log(f"[bench] FP32 vs FP64: {fp64}")
        except Exception as e:  # reported, never required
            log(f"[bench] fp64 leg failed: {e}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(scenes, args, ndyn)
            log(f"[bench] cpu reference: {cpu}")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

    s0 = scenes[0].settings
    strip = ("wall", "counts")
    out = {
        "metric": "body-steps/s & contacts/s at fixed residual; learned vs plain warm start",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "contacts_per_s": head["contacts_per_s"],
        "config": {"workload": name, "worlds": len(scenes), "bodies": sum(sc.nb_bodies for sc in scenes),
                   "dynamic_bodies": ndyn, "joints": nj, "dt": DT,
                   "velocity_solver": (f"residual {RESIDUAL} (cap {RESIDUAL_CAP})" if batched
                                       else f"{int(s0['vel_iters'])} iterations"),
                   "position_iterations": int(s0["pos_iters"]),
                   "warm_start": "learned (MLP, BF16 MFMA)" if batched else "plain (cached impulses)",
                   "step": "hipGraph replay" if graph else "eager",
                   "group_bodies": args.group_bodies or 256,
                   "pairs": c1["nb_pairs"], "manifolds": c1["nb_manifolds"],
                   "contact_points": c1["nb_contact_points"], "colors": c1["nb_colors"]},
        "learned": ({k: v for k, v in learned.items() if k not in strip} if learned else None),
        "plain": {k: v for k, v in plain.items() if k not in strip},
        "learned_vs_plain": ({"iterations_ratio": round(learned["mean_velocity_iterations"] /
                                                        max(1e-9, plain["mean_velocity_iterations"]), 4),
                              "throughput_ratio": round(learned["body_steps_per_s"] / plain["body_steps_per_s"], 4),
                              "same_launch_shape": tuned is not None,
                              "curves_from_initial_state": curves}
                             if learned else None),
        "training": training,
        "autotune": tuned,
        "fixed_iterations": fixed,
        "roofline": roofline,
        "kernels": kernels,
        "fp64_error": fp64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
