"""Writes the oracle's distances for the two benchmark workloads, so that the
GPU tests (and bench.py) check every device distance of those workloads
against the CPU restatement of X-TED_CPU without running it at full size:

  tests/golden/batch4096_dist.npy   int32[4096]: pair i = trees 2i, 2i+1 of
                                    bench.batch_trees(4096) (n ~ U[100, 1000],
                                    deep/shallow 50/50, alphabet 200)
  tests/golden/serving_dist.npz     the serving stream of bench.serve_bench
                                    (pool bench.serving_trees(10000), seed 7,
                                    50 queries/s for 10 s): qx, qy (pool
                                    indices) and dist per query

Pairs are spread over a process pool, one oracle thread each. Run from the
repository root (about 5 minutes on 8 cores):

    python tests/golden/make_workload_golden.py [--workers 8] [--only batch|serving]
"""
import argparse
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "x-ted_amd")]
import bench  # noqa: E402
import oracle  # noqa: E402
import xted  # noqa: E402

SERVE_SECONDS, SERVE_RATE, SERVE_SEED = 10.0, 50.0, 7

_trees = None


def _init(kind):
    global _trees
    _trees = bench.batch_trees(4096) if kind == "batch" else bench.serving_trees(10000)


def _dist(pair):
    i, j = pair
    d, _, _ = oracle.ted(_trees[i], _trees[j], threads=1, want_td=False)
    return int(d)


def _run(kind, pairs, workers):
    # Largest pairs first, so the pool's tail is short.
    trees = bench.batch_trees(4096) if kind == "batch" else bench.serving_trees(10000)
    S = np.array([xted.tree_stats(t[0])[0] for t in trees], np.int64)
    cells = np.array([S[i] * S[j] for i, j in pairs], np.int64)
    order = np.argsort(-cells)
    t = time.time()
    with Pool(workers, initializer=_init, initargs=(kind,)) as p:
        got = p.map(_dist, [pairs[k] for k in order], chunksize=1)
    out = np.empty(len(pairs), np.int32)
    out[order] = got
    print(f"{kind}: {len(pairs)} pairs, {cells.sum()} cells, {time.time() - t:.0f} s", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=os.cpu_count() or 4)
    ap.add_argument("--only", choices=["batch", "serving"])
    args = ap.parse_args()
    gold = os.path.join(ROOT, "tests", "golden")
    if args.only in (None, "batch"):
        d = _run("batch", [(2 * i, 2 * i + 1) for i in range(4096)], args.workers)
        np.save(os.path.join(gold, "batch4096_dist.npy"), d)
    if args.only in (None, "serving"):
        _, qx, qy = bench.serving_stream(10000, SERVE_SECONDS, SERVE_RATE, SERVE_SEED)
        d = _run("serving", list(zip(qx.tolist(), qy.tolist())), args.workers)
        np.savez_compressed(os.path.join(gold, "serving_dist.npz"), qx=qx.astype(np.int32),
                            qy=qy.astype(np.int32), dist=d,
                            stream=np.array([SERVE_SECONDS, SERVE_RATE, SERVE_SEED]))


if __name__ == "__main__":
    main()
