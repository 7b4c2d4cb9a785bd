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
  * the dominant kernel's roofline: 4 bytes of TD per cell against 8 TB/s
    HBM, timed with events around each of its launches;
  * the reference's own X-TED_CPU (compiled from its sources into
    oracle/_ref/) on a bounded sample of the same pairs, on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "x-ted_amd"))

import xted  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 4
CONFIG_BATCH = 3


def batch_trees(npairs, cfg=CONFIG_BATCH):
    """Trees 2i and 2i+1 form pair i."""
    trees = []
    for t in range(2 * npairs):
        s = xted.derive_seed(cfg, t)
        n = 100 + s % 901
        kind = xted.GEN_DEEP if (s >> 20) & 1 else xted.GEN_SHALLOW
        trees.append(xted.generate(kind, int(n), 200, s))
    return trees


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(trees, budget_s):
    """Reference X-TED_CPU on the first pairs of the workload until the time
    budget is spent (at least one pair)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.ref_available():
        return None
    threads = min(16, os.cpu_count() or 1)
    done, cells, t0 = 0, 0, time.perf_counter()
    while done < len(trees) // 2:
        x, y = trees[2 * done], trees[2 * done + 1]
        oracle.ref_ted(x, y, threads=threads, want_td=True)
        cells += xted.tree_stats(x[0])[0] * xted.tree_stats(y[0])[0]
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "gcups": cells / dt / 1e9,
            "sample": f"first {done} pairs of the 4096-pair batch, one pair at a time "
                      f"(X-TED_CPU, {threads} threads, {dt:.1f} s)"}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(pairs would shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    trees = batch_trees(args.pairs)
    log(f"generated {len(trees)} trees in {time.perf_counter() - t0:.1f} s")
    ctx = xted.Context(0)
    ids = ctx.pool_add(trees)
    xs, ys = ids[0::2].copy(), ids[1::2].copy()
    import autotune
    if args.autotune:
        rule, _ = autotune.tune(ctx, xs, ys, log=log)
    else:
        rule = autotune.load()
        autotune.apply(ctx, rule)
    log(f"granularity rule: {rule}")

    dist = None
    for _ in range(args.warmup):
        dist, _, st = ctx.batch(xs, ys)
        log(f"warmup done: {st.tables} tables, {st.cells} cells, {st.levels} levels")

    ctx.set_profile(True)
    times = []
    for k in range(args.steps):
        t = time.perf_counter()
        d, _, st = ctx.batch(xs, ys)
        times.append(time.perf_counter() - t)
        if dist is None:
            dist = d
        if not np.array_equal(d, dist):
            raise SystemExit("batched distances changed between steps")
        log(f"step {k}: {times[-1] * 1e3:.1f} ms (plan {st.plan_ms:.1f} ms, dp {st.dp_ms:.1f} ms)")
    prof = ctx.class_profile()
    ctx.set_profile(False)
    ms_per_step = 1e3 * sum(times) / len(times)
    pairs_per_s = args.pairs / (ms_per_step / 1e3)
    gcups = st.cells / (ms_per_step / 1e3) / 1e9

    # Dominant kernel: the class with the most device time in the last step.
    dom = max(prof.items(), key=lambda kv: kv[1][3])
    cls, (launches, tables, cells, ms) = dom
    achieved = BYTES_PER_CELL * cells / (ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": xted.class_name(cls), "launches": launches,
                "avg_launch_ms": round(ms / launches, 4),
                "bytes_per_launch": int(BYTES_PER_CELL * cells / launches)}
    classes = {xted.class_name(c): {"launches": v[0], "tables": v[1], "cells": v[2],
                                    "ms": round(v[3], 3)} for c, v in sorted(prof.items())}

    # Plain per-pair drop-in path on the first pairs of the batch.
    n_plain = min(args.plain_pairs, args.pairs)
    t = time.perf_counter()
    for i in range(n_plain):
        dp, _, _ = ctx.ted(trees[2 * i], trees[2 * i + 1], want_td=True)
        if dp != dist[i]:
            raise SystemExit(f"plain path distance differs from batch on pair {i}")
    plain_s = time.perf_counter() - t
    plain_pairs_per_s = n_plain / plain_s
    log(f"plain: {n_plain} pairs in {plain_s:.2f} s")
    ctx.close()

    cpu = None if args.no_cpu else cpu_baseline(trees, args.cpu_budget)

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
        "plain_pairs_per_s": round(plain_pairs_per_s, 2),
        "plain_sample_pairs": n_plain,
        "batched_vs_plain": round(pairs_per_s / plain_pairs_per_s, 2),
        "roofline": roofline,
        "classes": classes,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
