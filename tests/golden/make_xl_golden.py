"""Writes tests/golden/xl_expected.json: distance and two TD checksums of
pairs too large for one workgroup per table, computed by the oracle (the
CPU restatement of X-TED_CPU) on the seeded trees. The GPU tests rebuild
the same trees and compare. Run from the repository root:

    python tests/golden/make_xl_golden.py [--threads 6]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "x-ted_amd")]
import oracle  # noqa: E402
import xted  # noqa: E402

# name: (generator, nodes, alphabet, config id, tree indices)
CASES = {
    "rrt20000_a500": (xted.GEN_RRT, 20000, 500, 4, (0, 1)),
    "deep5000_a200": (xted.GEN_DEEP, 5000, 200, 5, (2, 3)),
    "rrt9000_a500": (xted.GEN_RRT, 9000, 500, 4, (4, 5)),
}


def trees_of(case):
    kind, n, alpha, cfg, (i, j) = CASES[case]
    return (xted.generate(kind, n, alpha, xted.derive_seed(cfg, i)),
            xted.generate(kind, n, alpha, xted.derive_seed(cfg, j)))


def td_checksums(td):
    """(sum of TD, sum of TD[i][j] * (((i*n + j) mod 1000003) + 1)) modulo 2^64."""
    m, n = td.shape
    s1 = np.uint64(0)
    s2 = np.uint64(0)
    cols = np.arange(n, dtype=np.uint64)
    for r0 in range(0, m, 512):
        blk = td[r0:r0 + 512].astype(np.uint64)
        rows = np.arange(r0, r0 + blk.shape[0], dtype=np.uint64)[:, None]
        w = (rows * np.uint64(n) + cols[None, :]) % np.uint64(1000003) + np.uint64(1)
        s1 = s1 + blk.sum(dtype=np.uint64)
        s2 = s2 + (blk * w).sum(dtype=np.uint64)
    return int(s1), int(s2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    out_path = os.path.join(ROOT, "tests", "golden", "xl_expected.json")
    out = {}
    if os.path.exists(out_path):
        with open(out_path) as f:
            out = json.load(f)
    for name in CASES:
        if args.only and name not in args.only:
            continue
        x, y = trees_of(name)
        t = time.time()
        d, td, _ = oracle.ted(x, y, threads=args.threads)
        s1, s2 = td_checksums(td)
        out[name] = {"dist": int(d), "td_sum": s1, "td_wsum": s2, "m": len(x[0]), "n": len(y[0])}
        print(name, out[name], f"{time.time() - t:.0f} s", flush=True)
        del td
        with open(out_path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
