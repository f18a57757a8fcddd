#!/usr/bin/env python3
"""Benchmark of the MI355X PhysicsWorld::update() step path.

Default workload (one MI355X): 65,536 mixed dynamic bodies (1/3 spheres r~U[0.25,1],
1/3 boxes half-extents ~U[0.25,1], 1/3 capsules r~U[0.2,0.5] h~U[0.5,2], density
1000, uniform in a 200 x 200 x 100 m volume, random orientations, v ~ N(0,1))
inside a static floor + 4 walls, dt = 1/60 s, FP32.  Inputs are resident in HBM
before the timed region.  A step is one full PhysicsWorld::update(): broadphase,
narrowphase, contact creation, coloured solver, integration, refit.

Prints one JSON line (see the repository's bench contract).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, vendor figure


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic (compulsory) HBM bytes per launch of the kernels that can dominate a step: every datum
# a launch must read or write, counted once.  `c` = counts of the step, `nlaunch` = launches per step.
def bytes_per_launch(name, c):
    nt = c["nb_tests"]
    m, p = c["nb_manifolds"], c["nb_contact_points"]
    ncol = max(1, c["nb_colors"])
    nb, ncl, npairs = c["nb_bodies"], c["nb_colliders"], c["nb_pairs"]
    np_types = {"np_sphere_sphere": 1, "np_sphere_capsule": 2, "np_capsule_capsule": 3, "np_sphere_poly": 4,
                "np_capsule_poly": 5, "np_poly_poly": 6}
    if name in np_types:
        # 176 B / test (pair record + 2 collider transforms/shapes + cache) + 40 B / raw contact kept
        tests = nt[np_types[name]]
        return 176 * tests + 40 * p * (tests / max(1, sum(nt)))
    if name == "sv_solve_color":
        # one sweep over all colours: 204 B of rows per manifold + 112 B per point +
        # 2 x 212 B of body state per manifold; a launch is one colour
        return (204 * m + 112 * p + 424 * m) / ncol
    if name == "sv_warmstart":
        return (204 * m + 112 * p + 424 * m) / ncol
    if name == "sv_prepare":
        return 420 * m + 160 * p
    if name == "bp_traverse":
        return 48 * ncl + 4 * ncl + 8 * npairs / 2
    if name == "bp_merge":
        return npairs * (8 + 8 + 240 * 2)
    if name in ("dyn_integrate_velocities",):
        return nb * (16 * 9 + 16 * 5)
    if name == "dyn_integrate_positions":
        return nb * (16 * 6 + 16 * 2)
    if name == "dyn_update_state":
        return nb * (16 * 5 + 16 * 7)
    if name == "bp_refit":
        return ncl * (16 * 6 + 16 * 6 + 8)
    return None


KERNELS = ["bp_bounds", "bp_morton", "bp_sort", "bp_karras", "bp_fit", "bp_world_roots", "bp_traverse", "bp_merge",
           "np_classify", "np_sphere_sphere", "np_sphere_capsule", "np_capsule_capsule", "np_sphere_poly",
           "np_capsule_poly", "np_poly_poly", "sv_flag", "sv_compact", "sv_adj_count", "sv_adj_fill",
           "sv_color_round", "sv_color_keys", "sv_color_ranges", "dyn_integrate_velocities", "sv_prepare",
           "sv_warmstart", "sv_solve_color", "sv_store", "dyn_integrate_positions", "dyn_update_state", "bp_refit"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    from reactphysics3d_amd import scene as S
    from reactphysics3d_amd.lib import Batch

    dt = 1.0 / 60.0
    t0 = time.time()
    sc = S.mixed_container(args.bodies, seed=2)
    log(f"[bench] scene {sc.name}: {sc.nb_bodies} bodies built in {time.time() - t0:.1f}s")
    t0 = time.time()
    dev = Batch(sc)
    log(f"[bench] uploaded in {time.time() - t0:.1f}s")
    for _ in range(args.warmup):
        dev.step(dt)
    dev.sync()
    log(f"[bench] warmup {args.warmup} steps done; counts {dev.counts()}")

    if not args.no_profile:
        dev.profile(True)
    contacts = 0
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.step(dt)
        contacts += dev.counts()["nb_contact_points"]
    dev.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    counts = dev.counts()
    ndyn = sc.nb_dynamic
    value = ndyn * args.steps / wall
    log(f"[bench] timed {args.steps} steps: {wall * 1e3 / args.steps:.3f} ms/step, {value:.4g} body-steps/s")

    roofline = None
    breakdown = {}
    if not args.no_profile:
        for k in KERNELS:
            ms, n = dev.kernel_time(k)
            if n:
                breakdown[k] = (ms, n)
        total = sum(ms * n for ms, n in breakdown.values())
        for k, (ms, n) in sorted(breakdown.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
            log(f"[bench]   {k:28s} {ms * 1e3:9.1f} us x {n / args.steps:6.1f}/step  "
                f"{100 * ms * n / max(total, 1e-12):5.1f}%")
        ranked = sorted(breakdown.items(), key=lambda kv: -kv[1][0] * kv[1][1])
        for name, (ms, n) in ranked:
            b = bytes_per_launch(name, counts)
            if b is None:
                continue
            achieved = b / (ms * 1e-3) / 1e9
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                        "bytes_per_launch": int(b), "mean_launch_us": round(ms * 1e3, 2),
                        "share_of_step": round(ms * n / max(total, 1e-12), 3)}
            break

    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref
            if ref.available():
                threads = args.cpu_threads or min(16, os.cpu_count() or 1)
                secs, cts = ref.cpu_bench(sc, threads, 0, args.cpu_steps, dt)
                cpu = {"value": round(ndyn * args.cpu_steps * threads / secs, 1), "unit": "body-steps/s",
                       "cores": threads, "kind": "reference",
                       "contacts_per_s": round(cts / secs, 1),
                       "sample": f"{threads} threads x 1 copy each of the same {sc.nb_bodies}-body world, "
                                 f"first {args.cpu_steps} steps, ReactPhysics3D 0.10.2 g++ -O3 FP32"}
                log(f"[bench] cpu reference: {cpu}")
            else:
                log("[bench] reference harness not built; no CPU baseline")
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")

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
        "contacts_per_s": round(contacts / wall, 1),
        "config": {"workload": f"mixed_container_{args.bodies}", "bodies": sc.nb_bodies, "dynamic_bodies": ndyn,
                   "dt": dt, "velocity_iterations": int(sc.settings["vel_iters"]),
                   "position_iterations": int(sc.settings["pos_iters"]), "worlds": 1,
                   "pairs": counts["nb_pairs"], "manifolds": counts["nb_manifolds"],
                   "contact_points": counts["nb_contact_points"], "colors": counts["nb_colors"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
