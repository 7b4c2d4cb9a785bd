"""Autotuner for the granularity cut-overs.

The reference fixes its granularity thresholds at 15 / 18,750 / 25,250,000
cells (X-TED_GPU/parallel_X_TED.cu:382) from a cost model fitted to one GPU
(supplement, section 2). Here the class rule has four knobs
(include/xted.h, xted_set_tuning): slot bytes per wave before a table gets
more waves, the most columns per lane, and a minimum wave count above a
column threshold. This module times a workload under a grid of rules on the
device it runs on and keeps the fastest. Every candidate must reproduce the
distances of the default rule exactly, or it is rejected.

    python x-ted_amd/autotune.py [--pairs N] [--out x-ted_amd/tuning.json]
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import xted  # noqa: E402

DEFAULT = {"lds_per_wave": 10 * 1024, "max_cpl": 8, "min_waves": 1, "wide_cols": 1 << 30,
           "tail_pairs": 32, "level_align": 0, "size_bands": 1}
# Searched in stages: the tail rule first, then the throughput rule with the
# best tail setting.
STAGES = [
    {"tail_pairs": [0, 8, 32, 128, 512]},
    {"lds_per_wave": [5 * 1024, 10 * 1024, 20 * 1024], "max_cpl": [4, 8]},
]
TUNING_FILE = os.path.join(HERE, "tuning.json")


def apply(ctx, rule):
    ctx.set_tuning(rule["lds_per_wave"], rule["max_cpl"], rule["min_waves"], rule["wide_cols"])
    ctx.set_tail_rule(rule["tail_pairs"])
    ctx.set_level_align(rule["level_align"])
    ctx.set_size_bands(rule["size_bands"])


def load(path=TUNING_FILE):
    """The stored rule, or the default when none was stored."""
    try:
        with open(path) as f:
            rule = json.load(f)["rule"]
        return {k: int(rule.get(k, DEFAULT[k])) for k in DEFAULT}
    except (OSError, KeyError, ValueError):
        return dict(DEFAULT)


def time_rule(ctx, xs, ys, rule, reps=2):
    apply(ctx, rule)
    dist, _, _ = ctx.batch(xs, ys)  # warm
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        d, _, _ = ctx.batch(xs, ys)
        best = min(best, time.perf_counter() - t)
        if not np.array_equal(d, dist):
            raise RuntimeError("distances changed between runs")
    return best, dist


def tune(ctx, xs, ys, stages=STAGES, log=print):
    """Returns (best_rule, results) over the staged grid; the default rule's
    distances are the reference every candidate must match."""
    base_t, base_d = time_rule(ctx, xs, ys, DEFAULT)
    results = [(dict(DEFAULT), base_t)]
    log(f"default {DEFAULT}: {base_t * 1e3:.1f} ms")
    best, best_t = dict(DEFAULT), base_t
    for grid in stages:
        keys = list(grid)
        start = dict(best)
        for vals in itertools.product(*(grid[k] for k in keys)):
            rule = dict(start, **dict(zip(keys, vals)))
            if any(rule == r for r, _ in results):
                continue
            try:
                t, d = time_rule(ctx, xs, ys, rule)
            except xted.XtedError as e:
                log(f"{rule}: rejected ({e})")
                continue
            if not np.array_equal(d, base_d):
                raise RuntimeError(f"rule {rule} changed a distance")
            results.append((rule, t))
            log(f"{rule}: {t * 1e3:.1f} ms")
            if t < best_t:
                best, best_t = rule, t
    apply(ctx, best)
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from bench import batch_trees
    trees = batch_trees(args.pairs)
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        best, results = tune(ctx, ids[0::2].copy(), ids[1::2].copy())
    out = {"rule": best, "workload": f"batch{args.pairs}_uneven_n100-1000_deep50_shallow50",
           "results_ms": [{"rule": r, "ms": round(t * 1e3, 2)} for r, t in results]}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"best": best}))


if __name__ == "__main__":
    main()
