"""Summarise rocprofv3 results (rocpd SQLite databases) as short markdown.

    python -m reactphysics3d_amd.profile_summary --trace DB [--fetch DB --write DB] [--copy-bytes N]
                                                 [--command "..."] [--top 25]

--trace  : database of a `rocprofv3 --kernel-trace --stats` run; prints per-kernel calls, mean and
           total duration and share of GPU time.
--fetch / --write : databases of `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs of the same
           command; prints the mean KiB per dispatch for every kernel.
--copy-bytes : size in bytes of the plain elementwise copy that the profiled command ran
           (`bench.py --copy-probe`): its measured FETCH/WRITE KiB divided by the known bytes gives
           the scale factor that turns the counters into HBM bytes.
--traffic-json : also write the scaled HBM bytes per dispatch of every kernel as JSON (bench.py
           reads profiles/traffic_<workload>.json for the roofline's "traffic").
--min-us KERNEL=US : count only the dispatches of KERNEL that ran at least US microseconds (repeatable;
           separates, e.g., residual-mode solver launches from fixed-iteration ones in one run).
--pmc DB : database of any `rocprofv3 --pmc ...` run (repeatable); prints, for every counter it holds,
           the mean value per dispatch of each kernel.  With --copy-bytes, a counter named in
           --bytes-read / --bytes-write (repeatable) is also turned into bytes by the value it reads on
           the copy kernel, which moves copy-bytes each way.
--kernels K1,K2 : restrict the --pmc tables to these kernels (and the copy kernel).
--top-dispatches KERNEL=N : count only the N dispatches of KERNEL with the largest value of each
           counter (repeatable; for counter runs whose records carry no usable dispatch times).
"""
from __future__ import annotations

import argparse
import re
import sqlite3
import sys
from collections import defaultdict


def short_name(name: str) -> str:
    """`rp3dg::np_poly_poly(rp3dg::NpArgs, ...)` -> `np_poly_poly`; torch kernels keep a short tag."""
    base = name.split("(")[0]
    base = re.sub(r"<.*", "", base)
    return base.split("::")[-1][:60]


def kernel_trace(db: str, min_us=None):
    """Per kernel: list of dispatch durations in microseconds (from start/end timestamps in ns)."""
    c = sqlite3.connect(db)
    rows = c.execute("select name, start, end from kernels").fetchall()
    agg = defaultdict(list)
    for name, t0, t1 in rows:
        k, us = short_name(name), (t1 - t0) / 1e3
        if min_us and us < min_us.get(k, 0.0):
            continue
        agg[k].append(us)
    return agg


TOP = {}   # kernel -> number of largest dispatches kept (--top-dispatches)


def _keep(k, vals):
    n = TOP.get(k)
    return sorted(vals)[-n:] if n else vals


def counter_means(db: str, counter: str, min_us=None):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value, start, end from counters_collection where counter_name = ?",
                     (counter,)).fetchall()
    agg = defaultdict(list)
    for name, v, t0, t1 in rows:
        k = short_name(name)
        if min_us and (t1 - t0) / 1e3 < min_us.get(k, 0.0):
            continue
        agg[k].append(float(v))
    agg = {k: _keep(k, v) for k, v in agg.items()}
    return {k: (sum(v) / len(v), len(v)) for k, v in agg.items() if v}


def all_counters(db: str, min_us=None):
    """{counter: {kernel: (mean per dispatch, dispatches)}} of every counter in a --pmc database."""
    c = sqlite3.connect(db)
    rows = c.execute("select counter_name, kernel_name, value, start, end from counters_collection").fetchall()
    agg = defaultdict(lambda: defaultdict(list))
    for cn, name, v, t0, t1 in rows:
        k = short_name(name)
        if min_us and (t1 - t0) / 1e3 < min_us.get(k, 0.0):
            continue
        agg[cn][k].append(float(v))
    out = {}
    for cn, ks in agg.items():
        ks = {k: _keep(k, v) for k, v in ks.items()}
        out[cn] = {k: (sum(v) / len(v), len(v)) for k, v in ks.items() if v}
    return out


def copy_kernel(means):
    for k in means:
        if "elementwise" in k or "vectorized" in k:
            return k
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--commit", default="", help="commit the measured tree was at (stamped into --traffic-json)")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--traffic-json", help="write {kernel: scaled HBM bytes per dispatch} to this file")
    ap.add_argument("--min-us", action="append", default=[], metavar="KERNEL=US")
    ap.add_argument("--pmc", action="append", default=[], metavar="DB")
    ap.add_argument("--bytes-read", action="append", default=[], metavar="COUNTER")
    ap.add_argument("--bytes-write", action="append", default=[], metavar="COUNTER")
    ap.add_argument("--kernels", default="")
    ap.add_argument("--top-dispatches", action="append", default=[], metavar="KERNEL=N")
    a = ap.parse_args(argv)
    min_us = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in a.min_us}
    TOP.update({kv.split("=")[0]: int(kv.split("=")[1]) for kv in a.top_dispatches})
    if min_us:
        out_filter = ", ".join(f"{k} >= {v:g} us" for k, v in min_us.items())
    out = []
    if a.command:
        out += ["Command:", "", "```", a.command, "```", ""]
    if min_us:
        out += [f"Dispatches counted: {out_filter} (other kernels: all).", ""]
    if TOP:
        out += ["Dispatches counted: " + ", ".join(f"the {n} largest of {k}" for k, n in TOP.items()) +
                " (other kernels: all).", ""]
    if a.trace:
        agg = kernel_trace(a.trace, min_us)
        tot = sum(sum(v) for v in agg.values()) or 1.0
        out += ["| kernel | calls | mean us | median us | min us | max us | total ms | share |",
                "|---|---:|---:|---:|---:|---:|---:|---:|"]
        for k, d in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[: a.top]:
            d = sorted(d)
            t = sum(d)
            out.append(f"| {k} | {len(d)} | {t / len(d):.1f} | {d[len(d) // 2]:.1f} | {d[0]:.1f} | {d[-1]:.1f} | "
                       f"{t / 1e3:.3f} | {100 * t / tot:.1f}% |")
        out.append("")
    if a.fetch or a.write:
        f = counter_means(a.fetch, "FETCH_SIZE", min_us) if a.fetch else {}
        w = counter_means(a.write, "WRITE_SIZE", min_us) if a.write else {}
        scale_f = scale_w = None
        ck = copy_kernel(f) or copy_kernel(w)
        if ck and a.copy_bytes > 0:
            if ck in f:
                scale_f = a.copy_bytes / (f[ck][0] * 1024.0)
            if ck in w:
                scale_w = a.copy_bytes / (w[ck][0] * 1024.0)
            out.append(f"Copy calibration: kernel `{ck}` moving {a.copy_bytes / 2**20:.0f} MiB each way; "
                       f"bytes per counted KiB: fetch {1024 * scale_f if scale_f else float('nan'):.1f}, "
                       f"write {1024 * scale_w if scale_w else float('nan'):.1f}")
            out.append("")
        out += ["| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB/dispatch (scaled) |",
                "|---|---:|---:|---:|---:|"]
        keys = sorted(set(f) | set(w), key=lambda k: -(f.get(k, (0, 0))[0] + w.get(k, (0, 0))[0]))
        for k in keys[: a.top]:
            fk = f.get(k, (float("nan"), 0))
            wk = w.get(k, (float("nan"), 0))
            hbm = ""
            if scale_f and scale_w:
                hbm = f"{(fk[0] * 1024 * scale_f + wk[0] * 1024 * scale_w) / 1e6:.3f}"
            out.append(f"| {k} | {max(fk[1], wk[1])} | {fk[0]:.1f} | {wk[0]:.1f} | {hbm} |")
        out.append("")
        if a.traffic_json and scale_f and scale_w:
            import json
            tj = {k: int(f[k][0] * 1024 * scale_f + w[k][0] * 1024 * scale_w) for k in set(f) & set(w)}
            tj["_measured"] = {"commit": a.commit, "command": a.command,
                               "fetch": a.fetch, "write": a.write, "copy_bytes": a.copy_bytes}
            with open(a.traffic_json, "w") as fh:
                json.dump(tj, fh, indent=1, sort_keys=True)
    for db in a.pmc:
        counters = all_counters(db, min_us)
        names = sorted(counters)
        kernels = set()
        for cn in names:
            kernels |= set(counters[cn])
        ck = copy_kernel({k: 0 for k in kernels})
        want = [k for k in a.kernels.split(",") if k]
        if want:
            kernels = {k for k in kernels if k in want or k == ck}
        first = counters[names[0]] if names else {}
        order = sorted(kernels, key=lambda k: -(first.get(k, (0.0, 0))[0]))
        out += [f"Counters of `{db}` (mean per dispatch):", ""]
        out += ["| kernel | dispatches | " + " | ".join(names) + " |",
                "|---|---:|" + "---:|" * len(names)]
        for k in order[: a.top]:
            disp = max(counters[cn].get(k, (0.0, 0))[1] for cn in names)
            vals = []
            for cn in names:
                v = counters[cn].get(k)
                vals.append(f"{v[0]:.6g}" if v else "")
            out.append(f"| {k} | {disp} | " + " | ".join(vals) + " |")
        out.append("")
        scaled = [cn for cn in names if cn in a.bytes_read or cn in a.bytes_write]
        if scaled and ck and a.copy_bytes > 0:
            out += [f"Bytes per dispatch, each counter scaled by its value on `{ck}` "
                    f"({a.copy_bytes / 2**20:.0f} MiB read and written):", ""]
            out += ["| kernel | " + " | ".join(f"{cn} MB" for cn in scaled) + " |",
                    "|---|" + "---:|" * len(scaled)]
            for k in order[: a.top]:
                vals = []
                for cn in scaled:
                    base = counters[cn].get(ck, (0.0, 0))[0]
                    v = counters[cn].get(k)
                    vals.append(f"{v[0] / base * a.copy_bytes / 1e6:.3f}" if v and base > 0 else "")
                out.append(f"| {k} | " + " | ".join(vals) + " |")
            out.append("")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
