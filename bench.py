#!/usr/bin/env python3
"""Benchmark: tree edit distance pairs/s on one MI355X.

Workload (one step): the 4096-pair uneven batch. Every tree has n ~ U[100,
1000] nodes, the deep/narrow and shallow/wide generators are mixed 50/50,
and labels are uniform over an alphabet of 200. All 4096 distances are
computed in one device plan: levels by launch, tables by granularity class.
The trees are resident in HBM before the timed region starts; each step
covers plan building, every forest-DP launch and the distance read-back.

The same JSON line also reports:
  * the plain per-pair drop-in path (xted_ted: upload, plan, DP, full-table
    read-back per pair) on a sample of the same pairs;
  * GCUPS (interior forest cells / s);
  * the dominant kernel's roofline: 4 bytes of TD per cell (one int32 entry
    of the tree-distance table) against 8 TB/s HBM, timed with events around
    each of its launches; the arena stores 16-bit entries, so the bytes it
    moves are half that (frac_at_arena_bytes beside it);
  * the reference's own X-TED_CPU (compiled from its sources into
    oracle/_ref/) on a bounded sample of the same pairs, on this host;
  * all-pairs mode: the distance matrix of the first 128 trees (8128 pairs);
  * the same batch under a non-unit cost model (per-label delete / insert
    costs and a rename matrix): pairs/s and its dominant kernel's roofline;
  * the extra-large pair (two 20,000-node random recursive trees): GCUPS
    and algorithmic HBM rate;
  * serving: a pool of 10,000 trees (50..5000 nodes, log-uniform), Poisson
    arrivals of random pairs at a fixed rate through the continuous-batching
    server, p50/p99 latency and completed pairs/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "x-ted_amd"))

import xted  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per interior cell: one int32 TD entry read (forest cells)
# or written (tree cells). The arena holds each entry in 16 bits.
BYTES_PER_CELL = 4
ARENA_BYTES_PER_CELL = 2
# Seed streams of the synthetic workloads (xted.derive_seed(stream, tree index)).
SEED_BATCH = 3
SEED_SERVE = 5
SEED_XL = 4


def batch_trees(npairs, cfg=SEED_BATCH):
    """Trees 2i and 2i+1 form pair i."""
    trees = []
    for t in range(2 * npairs):
        s = xted.derive_seed(cfg, t)
        n = 100 + s % 901
        kind = xted.GEN_DEEP if (s >> 20) & 1 else xted.GEN_SHALLOW
        trees.append(xted.generate(kind, int(n), 200, s))
    return trees


def serving_trees(ntrees, cfg=SEED_SERVE):
    """Pool for serving: n log-uniform in [50, 5000], deep/shallow 50/50,
    alphabet 200, seeded per tree."""
    trees = []
    for t in range(ntrees):
        s = xted.derive_seed(cfg, t)
        u = (s & 0xFFFFFFFF) / 2.0**32
        n = int(round(math.exp(math.log(50) + u * (math.log(5000) - math.log(50)))))
        kind = xted.GEN_DEEP if (s >> 40) & 1 else xted.GEN_SHALLOW
        trees.append(xted.generate(kind, n, 200, s))
    return trees


def serving_stream(ntrees, seconds, rate, seed=7):
    """Seeded Poisson query stream over a pool of ntrees trees: arrival
    times (s, sorted) and the x / y pool indices of uniformly random pairs."""
    rng = np.random.default_rng(seed)
    n_arr = max(1, int(rng.poisson(rate * seconds)))
    at = np.sort(rng.uniform(0.0, seconds, n_arr))
    qx = rng.integers(0, ntrees, n_arr)
    qy = rng.integers(0, ntrees, n_arr)
    return at, qx, qy


def serve_bench(trees, seconds, rate, budget, max_batch, seed=7):
    """Poisson arrivals of uniformly random pairs at `rate` per second for
    `seconds`, through the continuous-batching server. Returns the latency
    percentiles (submit -> result, measured by the server), the completed
    pairs/s and a check of a sample of results against one-pair batches."""
    at, qx, qy = serving_stream(len(trees), seconds, rate, seed)
    n_arr = len(at)
    ctx = xted.Context(0)
    t = time.perf_counter()
    ids = ctx.pool_add(trees)
    upload_s = time.perf_counter() - t
    srv = xted.Server(ctx, hbm_budget=budget, max_batch=max_batch)
    res = {}
    t0 = time.perf_counter()
    k = 0
    while len(res) < n_arr:
        now = time.perf_counter() - t0
        while k < n_arr and at[k] <= now:
            srv.submit(int(ids[qx[k]]), int(ids[qy[k]]), k)
            k += 1
        wait_ms = 1 if k < n_arr else 20
        for tag, status, dist, lat, queue in srv.poll(cap=4096, timeout_ms=wait_ms):
            res[tag] = (status, dist, lat, queue, time.perf_counter() - t0)
        if time.perf_counter() - t0 > seconds + 120:
            break
    srv.close()
    ok = [v for v in res.values() if v[0] == 0]
    lat = np.array([v[2] for v in ok])
    span = max(v[4] for v in res.values()) if res else seconds
    # Every served result against the oracle's distance when this is the
    # stream the committed fixture covers; otherwise a sample against
    # one-pair batches.
    gold = golden_serving(seconds, rate, seed)
    if gold is not None and len(gold) == n_arr and len(trees) == 10000:
        for tag, v in res.items():
            if v[0] == 0 and v[1] != gold[tag]:
                raise SystemExit(f"served distance differs from the oracle (query {tag})")
        checked = {"against": "oracle (tests/golden/serving_dist.npz)",
                   "queries": sum(1 for v in res.values() if v[0] == 0)}
    else:
        sample = sorted(res)[:: max(1, len(res) // 16)]
        for tag in sample:
            if res[tag][0] == 0:
                d = ctx.batch(np.array([ids[qx[tag]]]), np.array([ids[qy[tag]]]))[0][0]
                if d != res[tag][1]:
                    raise SystemExit(f"served distance differs from the one-pair batch "
                                     f"(query {tag})")
        checked = {"against": "one-pair batches", "queries": len(sample)}
    ctx.close()
    return {
        "offered_rate": rate, "seconds": seconds, "queries": n_arr,
        "completed": len(ok), "rejected": sum(1 for v in res.values() if v[0] != 0),
        "errors": {str(c): sum(1 for v in res.values() if v[0] == c)
                   for c in sorted({v[0] for v in res.values() if v[0] != 0})},
        "pairs_per_s": round(len(ok) / span, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 2) if len(lat) else None,
        "p99_ms": round(float(np.percentile(lat, 99)), 2) if len(lat) else None,
        "mean_queue_ms": round(float(np.mean([v[3] for v in ok])), 2) if ok else None,
        "hbm_budget": budget, "max_batch": max_batch, "pool_trees": len(trees),
        "pool_upload_s": round(upload_s, 2), "checked": checked,
    }


def golden_serving(seconds, rate, seed):
    """The oracle's distances for the serving stream (seconds, rate, seed)
    from tests/golden/serving_dist.npz, or None for another stream."""
    path = os.path.join(ROOT, "tests", "golden", "serving_dist.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    if [float(v) for v in g["stream"]] != [float(seconds), float(rate), float(seed)]:
        return None
    return g["dist"]


def golden_batch(npairs):
    """The oracle's distances for the 4096-pair batch
    (tests/golden/batch4096_dist.npy), or None for another size."""
    path = os.path.join(ROOT, "tests", "golden", "batch4096_dist.npy")
    if npairs != 4096 or not os.path.exists(path):
        return None
    return np.load(path)


def xl_pair_bench(reps=3):
    """The extra-large pair (two 20,000-node random recursive trees, alphabet
    500): its widest tables take the column-blocked form. Distance only
    (the 1.6 GB table stays in HBM); median of `reps` runs."""
    x = xted.generate(xted.GEN_RRT, 20000, 500, xted.derive_seed(SEED_XL, 0))
    y = xted.generate(xted.GEN_RRT, 20000, 500, xted.derive_seed(SEED_XL, 1))
    with xted.Context(0) as ctx:
        ids = ctx.pool_add([x, y])
        xs, ys = ids[:1].copy(), ids[1:].copy()
        ctx.batch(xs, ys)  # warm
        runs = []
        for _ in range(reps):
            t = time.perf_counter()
            d, _, st = ctx.batch(xs, ys)
            runs.append((time.perf_counter() - t, st, int(d[0])))
        # Then one run with every launch event-timed (per-class device
        # time, the root table's time per row); it is not the timed one.
        ctx.set_profile(True)
        _, _, pst = ctx.batch(xs, ys)
        prof = ctx.class_profile()
        launches = ctx.launch_profile()
        # And one with a single level stream (each launch alone): the
        # critical path of the pair's 38 levels.
        ctx.set_level_streams(1)
        ctx.batch(xs, ys)
        launches1 = ctx.launch_profile()
    wall, st, dist = sorted(runs, key=lambda r: r[0])[len(runs) // 2]
    xl = {xted.class_name(c): round(v[3], 2) for c, v in prof.items()
          if 27 <= c < 27 + len(xted.XL_BW)}
    classes = {xted.class_name(c): {"launches": v[0], "tables": v[1], "cells": v[2],
                                    "ms": round(v[3], 3)} for c, v in sorted(prof.items())}
    # The root table (the last level's one table, 20,000 x 20,000): its
    # device time per row is the latency one row of the column-blocked
    # pipeline costs.
    last = max(r["level"] for r in launches)
    root = max((r for r in launches if r["level"] == last), key=lambda r: r["ms"])
    with open(os.path.join(ROOT, "tests", "golden", "xl_expected.json")) as f:
        want = json.load(f)["rrt20000_a500"]["dist"]
    if dist != want:
        raise SystemExit(f"20k pair: distance {dist}, the oracle's is {want}")
    return {"pair": "rrt20000_a500", "dist": dist, "ms": round(wall * 1e3, 2),
            "gcups": round(st.cells / wall / 1e9, 2),
            "hbm_gbs_algorithmic": round(BYTES_PER_CELL * st.cells / wall / 1e9, 1),
            "frac_of_hbm_peak": round(BYTES_PER_CELL * st.cells / wall / 1e9 / HBM_PEAK_GBS, 4),
            "cells": st.cells, "tables": st.tables, "levels": st.levels,
            "plan_ms": round(st.plan_ms, 2), "dp_ms": round(st.dp_ms, 2),
            "profiled_ms": round(pst.total_ms, 2),
            "xl_stale_reads": st.xl_stale_reads, "xl_waits": st.xl_waits,
            "column_blocked_ms": xl, "classes": classes,
            "root_table": {"class": xted.class_name(root["cls"]), "ms": round(root["ms"], 3),
                           "rows": 20000, "us_per_row": round(root["ms"] * 1e3 / 20000, 3)},
            "level_profile": level_profile(launches1, launches),
            "checked": "distance against the oracle (tests/golden/xl_expected.json)"}


def row_latency(forms=((1000, (-1, 5, 9, 13, 17, 21, 29, 30, 31, 32)),
                        (4000, (-1, 28, 29, 30)))):
    """Time per row of one large table: the root table of a random recursive
    tree pair (alphabet 200), which runs alone in the last level, under each
    forced form (-1: the form the rule picks). One level stream; the best of
    three event-timed runs."""
    out = []
    for n, classes in forms:
        x = xted.generate(xted.GEN_RRT, n, 200, xted.derive_seed(901, n))
        y = xted.generate(xted.GEN_RRT, n, 200, xted.derive_seed(902, n))
        with xted.Context(0) as ctx:
            ids = ctx.pool_add([x, y])
            xs, ys = ids[:1].copy(), ids[1:].copy()
            ctx.set_level_streams(1)
            ref = None
            for cls in classes:
                ctx.force_class(cls)
                d, _, _ = ctx.batch(xs, ys)
                ref = int(d[0]) if ref is None else ref
                if int(d[0]) != ref:
                    raise SystemExit(f"row latency: forced class {cls} changed the distance")
                ctx.set_profile(True)
                best, root_cls = None, None
                for _ in range(3):
                    ctx.batch(xs, ys)
                    recs = ctx.launch_profile()
                    last = max(r["level"] for r in recs)
                    root = max((r for r in recs if r["level"] == last), key=lambda r: r["ms"])
                    if best is None or root["ms"] < best:
                        best, root_cls = root["ms"], root["cls"]
                ctx.set_profile(False)
                out.append({"n": n, "forced": cls, "form": xted.class_name(root_cls),
                            "root_ms": round(best, 3), "us_per_row": round(best * 1e3 / n, 3)})
                log(f"row latency {out[-1]}")
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Worker threads for the CPU baseline: the host's cores, capped at the
    CPU share a job gets beside one GPU (OMP_NUM_THREADS, 16 by default)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(os.cpu_count() or 1, share))


def cpu_baseline(trees, budget_s, threads):
    """Reference X-TED_CPU on the first pairs of the workload until the time
    budget is spent (at least one pair), one pair at a time as the
    reference's driver runs them (X-TED_CPU/main.cpp:7-10). Timed end to end
    (inputs -> distance and full table) and with the reference's own
    compute-phase timer (X-TED_CPU/parallel_computing.cpp:156-170,184)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.ref_available():
        return None
    done, cells, compute_ms, t0 = 0, 0, 0.0, time.perf_counter()
    while done < len(trees) // 2:
        x, y = trees[2 * done], trees[2 * done + 1]
        _, _, ms = oracle.ref_ted(x, y, threads=threads, want_td=True)
        compute_ms += ms
        cells += xted.tree_stats(x[0])[0] * xted.tree_stats(y[0])[0]
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "gcups": cells / dt / 1e9,
            "compute_phase_pairs_per_s": done / (compute_ms / 1e3),
            "compute_phase_gcups": cells / (compute_ms / 1e3) / 1e9,
            "ms_per_pair": 1e3 * dt / done, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"first {done} pairs of the 4096-pair batch, one pair at a time "
                      f"(X-TED_CPU, {threads} threads, {dt:.1f} s)"}


def copy_calibration(mib):
    """A plain copy of known size: rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE)
    on it gives the counter's KiB per copied KiB on this device."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst[-4:], src[-4:])
    print(json.dumps({"copy_bytes": 4 * n, "copies": 3}), flush=True)


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    profiles/traffic_<kernel>.txt (FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def dominant_roofline(prof, suffix=""):
    """Roofline of the class with the most device time in a profiled step
    (4 bytes of TD per interior cell over its event-timed launches, and the
    fraction at the arena's 2 bytes per cell) and the
    per-class table. suffix names the kernel variant in the profiles."""
    cls, (launches, tables, cells, ms) = max(prof.items(), key=lambda kv: kv[1][3])
    name = xted.class_name(cls) + suffix
    achieved = BYTES_PER_CELL * cells / (ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "frac_at_arena_bytes": round(achieved * ARENA_BYTES_PER_CELL / BYTES_PER_CELL
                                             / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profile(name),
                "kernel": name, "launches": launches,
                "avg_launch_ms": round(ms / launches, 4),
                "bytes_per_launch": int(BYTES_PER_CELL * cells / launches)}
    classes = {xted.class_name(c): {"launches": v[0], "tables": v[1], "cells": v[2],
                                    "ms": round(v[3], 3)} for c, v in sorted(prof.items())}
    return roofline, classes


def level_profile(log1, log4):
    """Per-level view of the event-timed steps (one launch record per forest
    launch, in launch order). With one level stream each launch runs alone,
    so a level's longest launch is the least time the level can take with
    its launches overlapped: the sum over levels is the batch's critical
    path under this plan (the latency bound), beside the time the levels
    actually take on 4 streams. Levels whose launches hold fewer workgroups
    than half the device's CUs (128 of 256) leave most CUs idle for their
    whole span."""
    def per_level(log):
        lv = {}
        for r in log:
            d = lv.setdefault(r["level"], {"start": r["start_ms"], "end": 0.0, "longest": 0.0,
                                           "tables": 0, "wgs": 0})
            d["start"] = min(d["start"], r["start_ms"])
            d["end"] = max(d["end"], r["start_ms"] + r["ms"])
            d["longest"] = max(d["longest"], r["ms"])
            d["tables"] += r["tables"]
            tiny = r["cls"] - xted.CLASS_TINY
            per_wg = 256 // xted.TINY_SEG[tiny] if 0 <= tiny < len(xted.TINY_SEG) else 1
            d["wgs"] += -(-r["tables"] // per_wg)
        return lv
    l1, l4 = per_level(log1), per_level(log4)
    span4 = {L: d["end"] - d["start"] for L, d in l4.items()}
    total4 = max(d["end"] for d in l4.values()) - min(d["start"] for d in l4.values())
    thin = [L for L, d in l4.items() if L >= 0 and d["wgs"] < 128]
    return {
        "levels": len([L for L in l4 if L >= 0]),
        "critical_path_ms": round(sum(d["longest"] for L, d in l1.items()), 2),
        "level_spans_ms_4_streams": round(sum(span4.values()), 2),
        "dp_span_ms_4_streams": round(total4, 2),
        "thin_levels": len(thin),
        "thin_level_ms_4_streams": round(sum(span4[L] for L in thin), 2),
        "thin_fraction": round(sum(span4[L] for L in thin) / total4, 4),
        "per_level_ms_4_streams": [round(span4[L], 3) for L in sorted(span4) if L >= 0],
        "per_level_longest_ms_1_stream": [round(l1[L]["longest"], 3) for L in sorted(l1)
                                          if L >= 0],
        "per_level_tables": [l4[L]["tables"] for L in sorted(l4) if L >= 0],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--plain-pairs", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the granularity rule first and use the fastest")
    ap.add_argument("--no-xl", action="store_true", help="skip the 20,000-node pair")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the levels one by one instead of replaying a captured graph")
    ap.add_argument("--profile-steps", type=int, default=1,
                    help="steps of the per-launch timed region (class times, roofline)")
    ap.add_argument("--batch-groups", type=int, default=0,
                    help="groups a batch is cut into (0: the library default)")
    ap.add_argument("--rule", default="",
                    help='JSON object overriding keys of the granularity rule, e.g. {"max_cpl": 16}')
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="also run the serving leg (Poisson arrivals) for this long")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-trees", type=int, default=10000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--xl-only", action="store_true",
                    help="only the 20,000-node pair (for profiling it alone)")
    ap.add_argument("--no-all-pairs", action="store_true",
                    help="skip the all-pairs matrix of 128 trees")
    ap.add_argument("--no-costs", action="store_true",
                    help="skip the batch under a non-unit cost model")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving leg (for profiling it alone)")
    ap.add_argument("--serve-rate2", type=float, default=150.0,
                    help="also serve a second stream at this offered rate (near saturation: "
                         "about 110-150 pairs/s complete on this pool; 0 skips it)")
    ap.add_argument("--serve-seconds2", type=float, default=5.0)
    ap.add_argument("--row-latency", action="store_true",
                    help="only the time per row of one large table under each forced form")
    ap.add_argument("--dump-launches", default="",
                    help="write the per-launch records of the event-timed steps (1 and 4 level "
                         "streams) to this JSON file")
    ap.add_argument("--profile-only", action="store_true",
                    help="only event-timed steps on one level stream (the run to trace with "
                         "rocprofv3 for the dominant kernel's per-dispatch duration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.row_latency:
        print(json.dumps({"row_latency": row_latency()}), flush=True)
        return
    if args.xl_only:
        print(json.dumps({"xl_pair": xl_pair_bench()}), flush=True)
        return
    if args.serve_only:
        strees = serving_trees(args.serve_trees)
        print(json.dumps({"serving": serve_bench(strees, args.serve_seconds, args.serve_rate,
                                                 budget=32 << 30, max_batch=4096)}), flush=True)
        if args.serve_rate2 > 0:
            print(json.dumps({"near_saturation": serve_bench(
                strees, args.serve_seconds2, args.serve_rate2, budget=32 << 30, max_batch=4096,
                seed=8)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(pairs would shard across GPUs by pair or by block of the all-pairs matrix, with no exchange)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    trees = batch_trees(args.pairs)
    log(f"generated {len(trees)} trees in {time.perf_counter() - t0:.1f} s")
    ctx = xted.Context(0)
    ids = ctx.pool_add(trees)
    xs, ys = ids[0::2].copy(), ids[1::2].copy()
    import autotune
    if args.autotune:
        rule, _, _ = autotune.tune(ctx, xs, ys, log=log)
    else:
        rule = autotune.load()
        if args.rule:
            rule.update({k: int(v) for k, v in json.loads(args.rule).items()})
        autotune.apply(ctx, rule)
    if args.batch_groups:
        ctx.set_batch_groups(args.batch_groups)
    ctx.set_graphs(not args.no_graphs and not args.profile_only)
    if args.profile_only:
        ctx.set_level_streams(1)
    log(f"granularity rule: {rule}")

    dist = None
    for _ in range(args.warmup):
        dist, _, st = ctx.batch(xs, ys)
        log(f"warmup done: {st.tables} tables, {st.cells} cells, {st.levels} levels")

    def timed(steps, profile, tag, expect):
        ctx.set_profile(profile)
        times, st = [], None
        for k in range(steps):
            t = time.perf_counter()
            d, _, st = ctx.batch(xs, ys)
            times.append(time.perf_counter() - t)
            if not np.array_equal(d, expect):
                raise SystemExit("batched distances changed between steps")
            log(f"{tag} {k}: {times[-1] * 1e3:.1f} ms (plan {st.plan_ms:.1f} ms, "
                f"dp {st.dp_ms:.1f} ms)")
        prof = ctx.class_profile() if profile else None
        ctx.set_profile(False)
        return times, prof, st

    if dist is None:
        dist, _, _ = ctx.batch(xs, ys)
    # Every distance against the oracle's (committed fixture).
    gold = golden_batch(args.pairs)
    checked_batch = None
    if gold is not None:
        bad = np.flatnonzero(dist != gold)
        if len(bad):
            raise SystemExit(f"{len(bad)} batched distances differ from the oracle, "
                             f"first pairs {bad[:8].tolist()}")
        checked_batch = {"against": "oracle (tests/golden/batch4096_dist.npy)",
                         "pairs": int(len(gold))}
    if args.profile_only:
        # Only event-timed steps on one level stream: the run rocprofv3
        # traces for the dominant kernel's per-dispatch duration.
        ctx.set_level_streams(1)
        _, prof1, _ = timed(max(1, args.profile_steps), True, "profiled step (1 stream)", dist)
        roof1, classes1 = dominant_roofline(prof1)
        print(json.dumps({"roofline": roof1, "classes_1_stream": classes1}), flush=True)
        return
    # The timed steps: levels replayed from a captured graph unless
    # --no-graphs. Then timed regions with every forest launch bracketed by
    # HIP events on its own stream (launched one by one): the per-class
    # device times and the dominant kernel's launch durations, once with the
    # launches of a level on one stream (each launch has the device to
    # itself: the kernel's own rate) and once on the 4 level streams the
    # timed steps use (launches of a level share the device).
    times, _, st = timed(args.steps, False, "step", dist)
    ctx.set_level_streams(1)
    ptimes1, prof1, _ = timed(max(1, args.profile_steps), True, "profiled step (1 stream)", dist)
    log1 = ctx.launch_profile()
    ctx.set_level_streams(4)
    ptimes, prof, _ = timed(max(1, args.profile_steps), True, "profiled step", dist)
    log4 = ctx.launch_profile()
    ms_per_step = 1e3 * sum(times) / len(times)
    pairs_per_s = args.pairs / (ms_per_step / 1e3)
    gcups = st.cells / (ms_per_step / 1e3) / 1e9
    roofline, classes1 = dominant_roofline(prof1)
    roofline4, classes = dominant_roofline(prof)
    levels = level_profile(log1, log4)
    if args.dump_launches:
        with open(args.dump_launches, "w") as f:
            json.dump({"one_stream": log1, "four_streams": log4}, f)

    # The same batch under a non-unit cost model: the row-scan forms with
    # per-label cost lookups and prefix sums (the closed-form leaf tables and
    # the tiny form are unit-cost only, so every table takes a row scan).
    cost_leg = None
    if not args.no_costs:
        rng = np.random.default_rng(5)
        dele = rng.integers(1, 9, 200).astype(np.int32)
        ins = rng.integers(1, 9, 200).astype(np.int32)
        ren = rng.integers(0, 17, (200, 200)).astype(np.int32)
        np.fill_diagonal(ren, 0)
        ctx.set_costs((dele, ins, ren))
        cdist, _, _ = ctx.batch(xs, ys)
        ctimes, _, cst = timed(args.steps, False, "cost step", cdist)
        _, cprof, _ = timed(1, True, "profiled cost step", cdist)
        ctx.set_costs(None)
        cms = 1e3 * sum(ctimes) / len(ctimes)
        croof, cclasses = dominant_roofline(cprof, "_cost")
        cost_leg = {"costs": "alphabet 200, delete and insert uniform in [1, 8], rename uniform "
                             "in [0, 16] (0 on equal labels), seed 5",
                    "pairs_per_s": round(args.pairs / (cms / 1e3), 1),
                    "ms_per_step": round(cms, 3), "gcups": round(cst.cells / cms / 1e6, 2),
                    "tables": cst.tables, "vs_unit_costs": round(ms_per_step / cms, 3),
                    "roofline": croof, "classes": cclasses}
        log(f"cost model: {cost_leg}")

    # All-pairs / dataset join: the distance matrix of the first 128 trees of
    # the workload (8128 pairs, in one plan); this clears the pool.
    all_pairs = None
    if not args.no_all_pairs:
        ap_trees = trees[:128]
        ctx.all_pairs(ap_trees)  # warm
        t = time.perf_counter()
        apd, apst = ctx.all_pairs(ap_trees)
        ap_s = time.perf_counter() - t
        if not (apd == apd.T).all():
            raise SystemExit("all-pairs matrix is not symmetric")
        npairs_ap = len(ap_trees) * (len(ap_trees) - 1) // 2
        all_pairs = {"trees": len(ap_trees), "pairs": npairs_ap, "ms": round(ap_s * 1e3, 2),
                     "pairs_per_s": round(npairs_ap / ap_s, 1),
                     "gcups": round(apst.cells / ap_s / 1e9, 2)}
        log(f"all pairs: {all_pairs}")

    # Plain per-pair drop-in path on the first pairs of the batch.
    n_plain = min(args.plain_pairs, args.pairs)
    t = time.perf_counter()
    for i in range(n_plain):
        dp, _, _ = ctx.ted(trees[2 * i], trees[2 * i + 1], want_td=True)
        if dp != dist[i]:
            raise SystemExit(f"plain path distance differs from batch on pair {i}")
    plain_s = time.perf_counter() - t
    plain_pairs_per_s = n_plain / plain_s if n_plain else None
    log(f"plain: {n_plain} pairs in {plain_s:.2f} s")
    ctx.close()

    cpu = None
    if not args.no_cpu:
        nthreads = cpu_threads()
        cpu = cpu_baseline(trees, args.cpu_budget, nthreads)
        if cpu is not None and (os.cpu_count() or 1) > nthreads:
            cpu["note"] = (f"{nthreads} of the host's {os.cpu_count()} cores: the CPU share of a "
                           f"job beside one GPU (OMP_NUM_THREADS)")

    xl = None
    if not args.no_xl:
        xl = xl_pair_bench()
        log(f"20k pair: {xl}")

    serving = None
    if args.serve_seconds > 0:
        t = time.perf_counter()
        strees = serving_trees(args.serve_trees)
        log(f"serving pool: {len(strees)} trees in {time.perf_counter() - t:.1f} s")
        serving = serve_bench(strees, args.serve_seconds, args.serve_rate,
                              budget=32 << 30, max_batch=4096)
        log(f"serving: {serving}")
        if args.serve_rate2 > 0:
            serving["near_saturation"] = serve_bench(strees, args.serve_seconds2, args.serve_rate2,
                                                     budget=32 << 30, max_batch=4096, seed=8)
            log(f"serving near saturation: {serving['near_saturation']}")

    out = {
        "metric": "TED pairs/s (4096-pair uneven batch, 100-1000 nodes, batched persistent "
                  "plan) + GCUPS; plain per-pair path beside it",
        "value": round(pairs_per_s, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generators, trees resident in HBM)",
        "config": {"workload": "batch4096_uneven_n100-1000_deep50_shallow50_alphabet200",
                   "pairs": args.pairs, "tables": st.tables, "cells": st.cells,
                   "levels": st.levels, "launches": st.launches, "granularity_rule": rule},
        "gcups": round(gcups, 2),
        "plan_ms": round(st.plan_ms, 3),
        "dp_ms": round(st.dp_ms, 3),
        "graphs": not args.no_graphs,
        "profiled_ms_per_step": round(1e3 * sum(ptimes) / len(ptimes), 3),
        "plain_pairs_per_s": round(plain_pairs_per_s, 2) if n_plain else None,
        "plain_sample_pairs": n_plain,
        "batched_vs_plain": round(pairs_per_s / plain_pairs_per_s, 2) if n_plain else None,
        "roofline": roofline,
        "roofline_4_streams": roofline4,
        "classes": classes,
        "classes_1_stream": classes1,
        "profiled_ms_per_step_1_stream": round(1e3 * sum(ptimes1) / len(ptimes1), 3),
        "level_profile": levels,
        "checked": checked_batch,
        "gcups_without_leaf_cells": round((st.cells - classes.get("leaf", {}).get("cells", 0))
                                          / (ms_per_step / 1e3) / 1e9, 2),
        "cpu_baseline": cpu,
        "cost_model": cost_leg,
        "all_pairs": all_pairs,
        "xl_pair": xl,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
