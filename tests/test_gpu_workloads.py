"""The benchmark workloads at their own size, against the oracle.

* The 4096-pair uneven batch (n ~ U[100, 1000] per tree, deep/shallow 50/50,
  alphabet 200, unit costs) run exactly as bench.py runs it: the stored
  granularity rule (x-ted_amd/tuning.json), graphs on, leaf and tiny forms.
  Every distance against the oracle's (tests/golden/batch4096_dist.npy,
  written by tests/golden/make_workload_golden.py), a sample of 64 pairs
  against the oracle run live here (both shapes, the 8 pairs with the most
  levels), and the full tables of 4 pairs.
* Serving from the 10,000-tree pool (50..5000 nodes, log-uniform): every
  query of bench.py's seeded Poisson stream (508 queries, the 5 largest pairs
  among them), replayed through the continuous-batching server, against the
  oracle's distances (tests/golden/serving_dist.npz).

The recurrence both follow is X-TED_CPU/parallel_computing.cpp:4-45."""
import os
import sys
import time

import numpy as np
import pytest

import oracle
import xted
from conftest import GOLDEN, ROOT

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import autotune  # noqa: E402
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch4096():
    trees = bench.batch_trees(4096)
    ctx = xted.Context(0)
    ids = ctx.pool_add(trees)
    autotune.apply(ctx, autotune.load())
    ctx.set_graphs(True)
    yield ctx, trees, ids[0::2].copy(), ids[1::2].copy()
    ctx.close()


def levels_of(tree_pair):
    """Levels of a pair: the keyroot nesting heights of both roots plus one."""
    (xp, _), (yp, _) = tree_pair
    kx = oracle.keyroot_depth(oracle.orl(xp), oracle.keyroots(oracle.orl(xp)))
    ky = oracle.keyroot_depth(oracle.orl(yp), oracle.keyroots(oracle.orl(yp)))
    return int(kx.max()) + int(ky.max()) + 1


def test_batch4096_tuned_rule_every_distance(batch4096):
    ctx, trees, xs, ys = batch4096
    expect = np.load(os.path.join(GOLDEN, "batch4096_dist.npy"))
    print("batch4096: capture", flush=True)
    d1, _, st = ctx.batch(xs, ys)
    print("batch4096: replay", flush=True)
    d2, _, _ = ctx.batch(xs, ys)  # replayed from the captured graph
    assert st.tables == 454413917 and st.levels == 116
    for d in (d1, d2):
        bad = np.flatnonzero(d != expect)
        assert len(bad) == 0, f"{len(bad)} distances differ, first pairs {bad[:8].tolist()}"


@pytest.mark.parametrize("align", [1, 2])
def test_batch4096_level_alignment(batch4096, align):
    """Pairs placed to end at the batch's last level (1) or midway (2)
    instead of starting at level 0: every distance against the oracle's."""
    ctx, trees, xs, ys = batch4096
    expect = np.load(os.path.join(GOLDEN, "batch4096_dist.npy"))
    ctx.set_level_align(align)
    try:
        print(f"batch4096: level alignment {align}", flush=True)
        d, _, st = ctx.batch(xs, ys)
    finally:
        ctx.set_level_align(autotune.load()["level_align"])
    assert st.levels == 116
    bad = np.flatnonzero(d != expect)
    assert len(bad) == 0, f"{len(bad)} distances differ, first pairs {bad[:8].tolist()}"


def test_batch4096_plan_order(batch4096):
    """Tables of each launch in plan order instead of largest first: every
    distance against the oracle's."""
    ctx, trees, xs, ys = batch4096
    expect = np.load(os.path.join(GOLDEN, "batch4096_dist.npy"))
    ctx.set_size_bands(False)
    try:
        print("batch4096: plan order", flush=True)
        d, _, _ = ctx.batch(xs, ys)
    finally:
        ctx.set_size_bands(autotune.load()["size_bands"])
    bad = np.flatnonzero(d != expect)
    assert len(bad) == 0, f"{len(bad)} distances differ, first pairs {bad[:8].tolist()}"


@pytest.mark.parametrize("light", [2048, 65536])
def test_batch4096_light_heavy_split(batch4096, light):
    """Size-band mode 2 with most row-scan tables heavy (2 KiB per wave) or
    none (64 KiB): every distance against the oracle's."""
    ctx, trees, xs, ys = batch4096
    expect = np.load(os.path.join(GOLDEN, "batch4096_dist.npy"))
    rule = autotune.load()
    ctx.set_size_bands(2)
    ctx.set_light_lds(light)
    try:
        print(f"batch4096: light/heavy split at {light} bytes per wave", flush=True)
        d, _, _ = ctx.batch(xs, ys)
    finally:
        ctx.set_size_bands(rule["size_bands"])
        ctx.set_light_lds(rule["light_lds"])
    bad = np.flatnonzero(d != expect)
    assert len(bad) == 0, f"{len(bad)} distances differ, first pairs {bad[:8].tolist()}"


@pytest.mark.parametrize("mask", [0, 1, 7])
def test_batch4096_tiny_forms(batch4096, mask):
    """The auto rule with no tiny form, the 16-lane one only, or all three
    (tables of up to 64 columns one per wave): every distance against the
    oracle's."""
    ctx, trees, xs, ys = batch4096
    expect = np.load(os.path.join(GOLDEN, "batch4096_dist.npy"))
    ctx.set_tiny_forms(mask)
    try:
        print(f"batch4096: tiny forms {mask}", flush=True)
        d, _, _ = ctx.batch(xs, ys)
    finally:
        ctx.set_tiny_forms(autotune.load()["tiny_forms"])
    bad = np.flatnonzero(d != expect)
    assert len(bad) == 0, f"{len(bad)} distances differ, first pairs {bad[:8].tolist()}"


def test_batch4096_sample_against_live_oracle(batch4096_sample):
    ctx, trees, xs, ys, sample, d = batch4096_sample
    for k in sample:
        od, _, _ = oracle.ted(trees[2 * k], trees[2 * k + 1], threads=16, want_td=False)
        assert d[k] == od, f"pair {k}"


@pytest.fixture(scope="module")
def batch4096_sample(batch4096):
    ctx, trees, xs, ys = batch4096
    lv = np.array([levels_of((trees[2 * k], trees[2 * k + 1])) for k in range(4096)])
    deepest = np.argsort(-lv, kind="stable")[:8]
    n = np.array([len(trees[2 * k][0]) for k in range(4096)])
    depth = np.array([xted.tree_stats(trees[2 * k][0])[1] for k in range(4096)])
    deep = np.flatnonzero(depth > 8)
    shallow = np.flatnonzero(depth <= 8)
    rng = np.random.default_rng(11)
    rest = np.concatenate([rng.choice(np.setdiff1d(deep, deepest), 28, replace=False),
                           rng.choice(np.setdiff1d(shallow, deepest), 28, replace=False)])
    sample = np.concatenate([deepest, rest])
    assert len(np.unique(sample)) == 64 and n[sample].min() >= 100
    d, _, _ = ctx.batch(xs, ys)
    return ctx, trees, xs, ys, sample, d


def test_batch4096_full_tables(batch4096, batch4096_sample):
    """The whole batch with tables requested for 4 of its pairs (the two with
    the most levels and two shallow ones): full TD against the oracle."""
    ctx, trees, xs, ys = batch4096
    sample = batch4096_sample[4]
    pick = [int(sample[0]), int(sample[1]), int(sample[-1]), int(sample[-2])]
    shapes = [None] * 4096
    for k in pick:
        shapes[k] = (len(trees[2 * k][0]), len(trees[2 * k + 1][0]))
    print(f"batch4096: tables of pairs {pick}", flush=True)
    d, tds, _ = ctx.batch(xs, ys, td_shapes=shapes)
    for k in pick:
        od, otd, _ = oracle.ted(trees[2 * k], trees[2 * k + 1], threads=16)
        assert d[k] == od
        mism = np.argwhere(tds[k] != otd)
        assert len(mism) == 0, f"pair {k}: {len(mism)} cells differ, first {mism[:5].tolist()}"


def test_serving_pool_10000_every_result():
    gold = np.load(os.path.join(GOLDEN, "serving_dist.npz"))
    seconds, rate, seed = (float(v) for v in gold["stream"])
    at, qx, qy = bench.serving_stream(10000, seconds, rate, int(seed))
    assert (qx == gold["qx"]).all() and (qy == gold["qy"]).all()
    trees = bench.serving_trees(10000)
    S = np.array([xted.tree_stats(t[0])[0] for t in trees], np.int64)
    largest = np.argsort(-(S[qx] * S[qy]))[:5]
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        srv = xted.Server(ctx, hbm_budget=32 << 30, max_batch=4096)
        # The stream replayed five times faster than recorded (about 250
        # queries/s), so that groups overlap on the device.
        t0 = time.perf_counter()
        got, k = {}, 0
        while len(got) < len(at) and time.perf_counter() - t0 < 120:
            now = (time.perf_counter() - t0) * 5.0
            while k < len(at) and at[k] <= now:
                srv.submit(int(ids[qx[k]]), int(ids[qy[k]]), k)
                k += 1
            for tag, status, dist, lat, queue in srv.poll(cap=1024, timeout_ms=1):
                got[tag] = (status, dist)
        srv.close()
    print(f"serving: {len(got)} of {len(at)} results", flush=True)
    assert len(got) == len(at)
    assert all(got[k][0] == xted.OK for k in got)
    served = np.array([got[k][1] for k in range(len(at))])
    bad = np.flatnonzero(served != gold["dist"])
    assert len(bad) == 0, f"{len(bad)} served distances differ, first queries {bad[:8].tolist()}"
    assert (served[largest] == gold["dist"][largest]).all()
