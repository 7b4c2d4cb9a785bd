"""Serving: queries that arrive while others are on the device join later
batches; every result must equal the same pair evaluated alone, and the
oracle. Budget admission rejects what can never fit."""
import random
import time

import numpy as np
import pytest

import oracle
import xted

pytestmark = pytest.mark.gpu


def _pool(n, seed, lo=2, hi=400):
    rng = np.random.default_rng(seed)
    return [xted.generate(int(rng.integers(0, 3)), int(rng.integers(lo, hi)), 25,
                          xted.derive_seed(90 + seed, i)) for i in range(n)]


def _drain(srv, want, timeout_s=120):
    got = {}
    t_end = time.time() + timeout_s
    while len(got) < want and time.time() < t_end:
        for tag, status, dist, lat, queue in srv.poll(timeout_ms=50):
            got[tag] = (status, dist, lat, queue)
    assert len(got) == want, f"only {len(got)} of {want} results"
    return got


def test_server_results_same_alone_and_in_the_mix():
    trees = _pool(40, 1)
    rnd = random.Random(5)
    queries = [(rnd.randrange(40), rnd.randrange(40)) for _ in range(160)]
    # Alone: each pair as a one-pair batch on its own context.
    with xted.Context(0) as solo:
        ids = solo.pool_add(trees)
        alone = [int(solo.batch(np.array([ids[a]]), np.array([ids[b]]))[0][0]) for a, b in queries]
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        srv = xted.Server(ctx, hbm_budget=1 << 30, max_batch=48)
        try:
            # Arrivals spread out, so queries land while earlier batches run.
            for k, (a, b) in enumerate(queries):
                srv.submit(int(ids[a]), int(ids[b]), k)
                if k % 7 == 0:
                    time.sleep(rnd.random() * 0.004)
            got = _drain(srv, len(queries))
        finally:
            srv.close()
    for k, (a, b) in enumerate(queries):
        status, dist, lat, queue = got[k]
        assert status == 0
        assert dist == alone[k], (k, a, b)
        assert lat >= queue >= 0
    for k in range(0, len(queries), 16):
        a, b = queries[k]
        assert got[k][1] == oracle.ted(trees[a], trees[b], threads=4, want_td=False)[0]


def test_server_budget_rejects_oversized():
    trees = _pool(6, 2, lo=300, hi=400) + _pool(6, 3, lo=5, hi=20)
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        # 8*m*n bytes per query: a 300x300 pair needs 720 KB, a 20x20 one 3.2 KB.
        srv = xted.Server(ctx, hbm_budget=100_000, max_batch=16)
        try:
            srv.submit(int(ids[0]), int(ids[1]), 0)   # too large for the budget
            srv.submit(int(ids[6]), int(ids[7]), 1)   # fits
            srv.submit(int(ids[8]), int(ids[9]), 2)
            got = _drain(srv, 3)
        finally:
            srv.close()
    assert got[0][0] == xted.ERR_TOO_LARGE and got[0][1] == -1
    for k, (a, b) in ((1, (6, 7)), (2, (8, 9))):
        assert got[k][0] == 0
        assert got[k][1] == oracle.ted(trees[a], trees[b], want_td=False)[0]


def test_server_rejects_unknown_tree():
    with xted.Context(0) as ctx:
        ctx.pool_add(_pool(2, 4))
        srv = xted.Server(ctx, hbm_budget=1 << 20)
        try:
            with pytest.raises(xted.XtedError):
                srv.submit(0, 99, 0)
        finally:
            srv.close()
