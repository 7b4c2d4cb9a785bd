"""GPU parity under non-unit edit costs (per-label delete / insert tables and
a rename matrix; the reference hard-codes unit costs,
X-TED_CPU/TED.cpp:111,115,122): the device's cost mode against the oracle's
cost restatement, bit-exact on distances and full tree-distance tables, over
random cost models (zero costs, asymmetric insert / delete, rename above
delete + insert), both table orientations, every row-scan class and every
column-blocked form forced, batches, and values near the 16-bit slot limit."""
import numpy as np
import pytest

import oracle
import xted

pytestmark = pytest.mark.gpu


def gen(kind, n, alphabet, cfg, idx, band=(0, 0)):
    return xted.generate(kind, n, alphabet, xted.derive_seed(cfg, idx), *band)


def gen_any(kind, n, alphabet, cfg, idx):
    """Deep trees get a depth band the node count can reach."""
    if kind == xted.GEN_DEEP and n < 8:
        kind = xted.GEN_RRT
    band = (n // 8, n // 3) if kind == xted.GEN_DEEP else (0, 0)
    return gen(kind, n, alphabet, cfg, idx, band)


def rand_costs(rng, alphabet, hi=9, zeros=False):
    lo = 0 if zeros else 1
    dele = rng.integers(lo, hi, alphabet).astype(np.int32)
    ins = rng.integers(lo, hi, alphabet).astype(np.int32)
    ren = rng.integers(0, 2 * hi, (alphabet, alphabet)).astype(np.int32)
    np.fill_diagonal(ren, 0)
    return dele, ins, ren


def unit_costs(alphabet):
    ren = np.ones((alphabet, alphabet), np.int32)
    np.fill_diagonal(ren, 0)
    return np.ones(alphabet, np.int32), np.ones(alphabet, np.int32), ren


@pytest.fixture
def cost_ctx(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_costs(None)
    gpu_ctx.force_class(xted.CLASS_AUTO, lane_enable=True)


def assert_same(ctx, x, y, costs, threads=8):
    d, td, _ = ctx.ted(x, y)
    od, otd = oracle.ted_costs(x, y, costs, threads=threads)
    assert d == od
    mism = np.argwhere(td != otd)
    assert len(mism) == 0, f"{len(mism)} TD cells differ, first {mism[:5].tolist()}"
    return d


@pytest.mark.parametrize("seed", range(8))
def test_random_cost_models(cost_ctx, seed):
    rng = np.random.default_rng(100 + seed)
    alphabet = int(rng.choice([3, 8, 40]))
    costs = rand_costs(rng, alphabet, hi=int(rng.choice([3, 9, 60])), zeros=seed % 2 == 1)
    cost_ctx.set_costs(costs)
    for i in range(4):
        kx, ky = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        n1, n2 = int(rng.integers(1, 260)), int(rng.integers(1, 260))
        x = gen_any(kx, n1, alphabet, 200 + seed, 2 * i)
        y = gen_any(ky, n2, alphabet, 200 + seed, 2 * i + 1)
        assert_same(cost_ctx, x, y, costs)
        assert_same(cost_ctx, y, x, costs)


def test_unit_tables_match_unit_mode(cost_ctx):
    """A cost model of unit tables takes the cost kernels and must reproduce
    unit mode, which takes the closed-form leaf tables as well."""
    x = gen(xted.GEN_DEEP, 1000, 30, 210, 0, (60, 120))
    y = gen(xted.GEN_SHALLOW, 900, 30, 210, 1)
    d0, t0, _ = cost_ctx.ted(x, y)
    cost_ctx.set_costs(unit_costs(30))
    d1, t1, _ = cost_ctx.ted(x, y)
    assert d0 == d1 and np.array_equal(t0, t1)


@pytest.mark.parametrize("kind", [xted.GEN_RRT, xted.GEN_DEEP, xted.GEN_SHALLOW])
def test_thousand_node_pairs(cost_ctx, kind):
    rng = np.random.default_rng(220 + kind)
    costs = rand_costs(rng, 50)
    cost_ctx.set_costs(costs)
    band = (60, 120) if kind == xted.GEN_DEEP else (0, 0)
    x = gen(kind, 1000, 50, 220, 2 * kind, band)
    y = gen(kind, 700, 50, 220, 2 * kind + 1, band)
    assert_same(cost_ctx, x, y, costs, threads=16)
    assert_same(cost_ctx, y, x, costs, threads=16)


@pytest.mark.parametrize("cls", list(range(1, 26)))
def test_every_row_class_forced(cost_ctx, cls):
    rng = np.random.default_rng(300 + cls)
    costs = rand_costs(rng, 6, zeros=cls % 3 == 0)
    cost_ctx.set_costs(costs)
    cost_ctx.force_class(cls, lane_enable=False)
    for i, (kx, ky, n1, n2) in enumerate([(0, 0, 90, 70), (1, 2, 120, 150), (2, 1, 80, 130)]):
        x = gen(kx, n1, 6, 300 + cls, 2 * i)
        y = gen(ky, n2, 6, 300 + cls, 2 * i + 1)
        assert_same(cost_ctx, x, y, costs)


@pytest.mark.parametrize("cls", [27, 28, 29, 30, 31, 32])
def test_every_column_blocked_form_forced(cost_ctx, cls):
    rng = np.random.default_rng(400 + cls)
    costs = rand_costs(rng, 12)
    cost_ctx.set_costs(costs)
    cost_ctx.force_class(cls)
    for i, (kx, ky, n1, n2) in enumerate([(0, 1, 300, 420), (1, 2, 900, 1200)]):
        x = gen(kx, n1, 12, 400 + cls, 2 * i)
        y = gen(ky, n2, 12, 400 + cls, 2 * i + 1)
        assert_same(cost_ctx, x, y, costs, threads=16)


@pytest.mark.parametrize("form", range(3))
@pytest.mark.parametrize("seed", range(4))
def test_tiny_class_forced(cost_ctx, seed, form):
    """Every table of at most 16, 32 or 64 columns in the cost variant of
    tiny form 0, 1 or 2 (four, two or one table per wave, one per 16-, 32- or
    64-lane segment), including the one-row tables of leaf keyroots, which
    the cost model plans like any other."""
    rng = np.random.default_rng(500 + seed)
    costs = rand_costs(rng, 5, zeros=seed % 2 == 1)
    cost_ctx.set_costs(costs)
    cost_ctx.force_class(xted.CLASS_TINY + form, lane_enable=False)
    seg = xted.TINY_SEG[form]
    for i, (kx, ky, n1, n2) in enumerate([(0, 2, 60, 50), (1, 1, 90, 120), (2, 0, seg, seg),
                                          (0, 0, 1, 30), (1, 2, seg + 1, seg - 1)]):
        x = gen_any(kx, n1, 5, 500 + seed, 2 * i)
        y = gen_any(ky, n2, 5, 500 + seed, 2 * i + 1)
        assert_same(cost_ctx, x, y, costs)
        assert_same(cost_ctx, y, x, costs)


def test_batch_against_oracle(cost_ctx):
    rng = np.random.default_rng(7)
    costs = rand_costs(rng, 20)
    trees = [gen_any(int(rng.integers(0, 3)), int(rng.integers(1, 500)), 20, 600, i)
             for i in range(30)]
    cost_ctx.pool_clear()
    ids = cost_ctx.pool_add(trees)
    cost_ctx.set_costs(costs)
    xs = rng.integers(0, 30, 40).astype(np.int32)
    ys = rng.integers(0, 30, 40).astype(np.int32)
    shapes = [(len(trees[a][0]), len(trees[b][0])) for a, b in zip(xs, ys)]
    dist, tds, _ = cost_ctx.batch(ids[xs], ids[ys], td_shapes=shapes)
    # trees added after the cost model is set get their prefixes too
    more = [gen(xted.GEN_RRT, 300, 20, 601, i) for i in range(2)]
    mid = cost_ctx.pool_add(more)
    d2, t2, _ = cost_ctx.batch(mid[:1], mid[1:], td_shapes=[(300, 300)])
    for k, (a, b) in enumerate(zip(xs, ys)):
        od, otd = oracle.ted_costs(trees[a], trees[b], costs, threads=4)
        assert dist[k] == od and np.array_equal(tds[k], otd)
    od, otd = oracle.ted_costs(more[0], more[1], costs, threads=4)
    assert d2[0] == od and np.array_equal(t2[0], otd)


def test_switching_cost_models(cost_ctx):
    rng = np.random.default_rng(9)
    x = gen(xted.GEN_RRT, 200, 10, 700, 0)
    y = gen(xted.GEN_DEEP, 180, 10, 700, 1, (20, 40))
    c1, c2 = rand_costs(rng, 10), rand_costs(rng, 10, hi=30)
    cost_ctx.set_costs(c1)
    assert_same(cost_ctx, x, y, c1)
    cost_ctx.set_costs(c2)
    assert_same(cost_ctx, x, y, c2)
    cost_ctx.set_costs(None)
    d, td, _ = cost_ctx.ted(x, y)
    od, otd, _ = oracle.ted(x, y, threads=8)
    assert d == od and np.array_equal(td, otd)


def test_values_near_the_slot_limit(cost_ctx):
    """Forest values up to 65535 fit the 16-bit row slots."""
    alphabet = 4
    dele = np.array([8000, 7000, 6000, 5000], np.int32)
    ins = np.array([7500, 6500, 5500, 4500], np.int32)
    ren = np.array([[0, 9000, 15000, 32767], [100, 0, 20000, 3],
                    [13000, 1, 0, 25000], [32767, 40, 11111, 0]], np.int32)
    costs = (dele, ins, ren)
    cost_ctx.set_costs(costs)
    rng = np.random.default_rng(11)
    for i in range(12):
        m, n = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        x = gen_any(int(rng.integers(0, 3)), m, alphabet, 800, 2 * i)
        y = gen_any(int(rng.integers(0, 3)), n, alphabet, 800, 2 * i + 1)
        if dele[x[1]].sum() + ins[y[1]].sum() > 65535:
            with pytest.raises(xted.XtedError):
                cost_ctx.ted(x, y)
            continue
        assert_same(cost_ctx, x, y, costs)


def test_empty_trees_and_rejections(cost_ctx):
    rng = np.random.default_rng(13)
    costs = rand_costs(rng, 6)
    cost_ctx.set_costs(costs)
    x = gen(xted.GEN_RRT, 30, 6, 900, 0)
    empty = (np.empty(0, np.int32), np.empty(0, np.int32))
    assert cost_ctx.ted(x, empty)[0] == int(costs[0][x[1]].sum())
    assert cost_ctx.ted(empty, x)[0] == int(costs[1][x[1]].sum())
    bad = (x[0], np.full(30, 6, np.int32))
    with pytest.raises(xted.XtedError):
        cost_ctx.ted(x, bad)
    with pytest.raises(xted.XtedError):
        cost_ctx.edit_mapping(x, x)
    with pytest.raises(xted.XtedError):
        cost_ctx.set_costs((costs[0], costs[1], np.full((6, 6), 40000, np.int32)))
    big = (np.full(6, 30000, np.int32), np.full(6, 30000, np.int32), costs[2])
    cost_ctx.set_costs(big)
    with pytest.raises(xted.XtedError):
        cost_ctx.ted(x, x)
    cost_ctx.set_costs(None)
    assert cost_ctx.ted(x, bad)[0] >= 0


def test_pool_trees_outside_the_alphabet(cost_ctx):
    """Trees added before the cost model may carry labels it does not cover:
    pairs of them are rejected, other pairs run."""
    wide = [gen(xted.GEN_RRT, 80, 40, 950, i) for i in range(2)]
    narrow = [gen(xted.GEN_SHALLOW, 70, 3, 951, i) for i in range(2)]
    cost_ctx.pool_clear()
    ids = cost_ctx.pool_add(wide + narrow)
    costs = rand_costs(np.random.default_rng(17), 3)
    cost_ctx.set_costs(costs)
    with pytest.raises(xted.XtedError):
        cost_ctx.batch(ids[:1], ids[2:3])
    with pytest.raises(xted.XtedError):
        cost_ctx.pool_add(wide[:1])
    dist, tds, _ = cost_ctx.batch(ids[2:3], ids[3:4], td_shapes=[(70, 70)])
    od, otd = oracle.ted_costs(narrow[0], narrow[1], costs, threads=4)
    assert dist[0] == od and np.array_equal(tds[0], otd)


def test_full_batch_under_costs_sampled(cost_ctx):
    """The 4096-pair uneven batch (100..1000 nodes) under a cost model: every
    table takes a row scan, so the first level's launches hold tens of
    millions of workgroups and go in pieces. A sample of pairs is checked
    against the oracle."""
    import sys
    from conftest import ROOT
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    trees = bench.batch_trees(4096)
    cost_ctx.pool_clear()
    ids = cost_ctx.pool_add(trees)
    rng = np.random.default_rng(5)
    costs = rand_costs(rng, 200)
    cost_ctx.set_costs(costs)
    dist, _, st = cost_ctx.batch(ids[0::2], ids[1::2])
    assert st.tables > 1 << 28
    for k in rng.choice(4096, 6, replace=False):
        od, _ = oracle.ted_costs(trees[2 * k], trees[2 * k + 1], costs, threads=16, want_td=False)
        assert dist[k] == od


def _drain(srv, want, timeout_s=120):
    import time
    got = {}
    t_end = time.time() + timeout_s
    while len(got) < want and time.time() < t_end:
        for tag, status, dist, lat, queue in srv.poll(timeout_ms=50):
            got[tag] = (status, dist)
    assert len(got) == want, f"only {len(got)} of {want} results"
    return got


def test_server_across_a_cost_model_change():
    """Serving under a cost model, and a change of model while queries are in
    flight: each of those finishes under the model it was planned with or
    fails with XTED_ERR_INVALID_ARG (never a mixed result); queries after the
    change use the new model."""
    rng = np.random.default_rng(21)
    trees = [gen_any(int(rng.integers(0, 3)), int(rng.integers(2, 300)), 10, 960, i)
             for i in range(30)]
    c1, c2 = rand_costs(rng, 10), rand_costs(rng, 10, hi=20)
    qs = [(int(rng.integers(30)), int(rng.integers(30))) for _ in range(60)]
    later = [(int(rng.integers(30)), int(rng.integers(30))) for _ in range(20)]
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        ctx.set_costs(c1)
        srv = xted.Server(ctx, hbm_budget=1 << 30, max_batch=16)
        try:
            for k, (a, b) in enumerate(qs):
                srv.submit(int(ids[a]), int(ids[b]), k)
            ctx.set_costs(c2)
            first = _drain(srv, len(qs))
            for k, (a, b) in enumerate(later):
                srv.submit(int(ids[a]), int(ids[b]), 100 + k)
            second = _drain(srv, len(later))
        finally:
            srv.close()
    for k, (a, b) in enumerate(qs):
        status, dist = first[k]
        assert status in (xted.OK, xted.ERR_INVALID_ARG), status
        if status == xted.OK:
            d1 = oracle.ted_costs(trees[a], trees[b], c1, threads=4, want_td=False)[0]
            d2 = oracle.ted_costs(trees[a], trees[b], c2, threads=4, want_td=False)[0]
            assert dist in (d1, d2), (k, dist, d1, d2)
    for k, (a, b) in enumerate(later):
        status, dist = second[100 + k]
        assert status == xted.OK
        assert dist == oracle.ted_costs(trees[a], trees[b], c2, threads=4, want_td=False)[0]


def comb_tree(depth, alphabet, seed):
    """Each spine node's first child the next spine node, its second a leaf
    (preorder: the whole deeper spine comes before the leaf): a right-to-left
    sweep holds every such leaf's value until it reaches its previous
    sibling, so the cost form of the leaf lines needs a stack `depth` deep."""
    parent = [-1]
    spine = [0]
    for _ in range(depth):
        parent.append(spine[-1])
        spine.append(len(parent) - 1)
    # the leaves, second children, in preorder after each spine node's subtree
    for node in reversed(spine[:-1]):
        parent.append(node)
    # renumber into preorder: children of a spine node are (next spine, leaf)
    kids = {}
    for c, pa in enumerate(parent[1:], 1):
        kids.setdefault(pa, []).append(c)
    order, st = [], [0]
    while st:
        v = st.pop()
        order.append(v)
        st.extend(reversed(kids.get(v, [])))
    nid = {v: i for i, v in enumerate(order)}
    pre = np.full(len(parent), -1, np.int32)
    for v in range(1, len(parent)):
        pre[nid[v]] = nid[parent[v]]
    rng = np.random.default_rng(seed)
    return pre, rng.integers(0, alphabet, len(parent)).astype(np.int32)


@pytest.mark.parametrize("depth", [40, 127, 128, 129, 200])
def test_cost_leaf_lines_stack_limit(cost_ctx, depth):
    """Pairs whose trees need a line stack up to the limit (128) take the
    cost form of the leaf kernel; deeper ones plan their leaf tables as row
    scans. Full tables against the oracle on both sides of the limit, with
    the other tree random."""
    rng = np.random.default_rng(depth)
    costs = rand_costs(rng, 9)
    cost_ctx.set_costs(costs)
    x = comb_tree(depth, 9, depth)
    y = gen(xted.GEN_RRT, 150, 9, 931, depth)
    assert_same(cost_ctx, x, y, costs)
    assert_same(cost_ctx, y, x, costs)


@pytest.mark.parametrize("seed", range(4))
def test_cost_leaf_form_on_and_off(cost_ctx, seed):
    """The same batch with the leaf lines in the cost form of the closed form
    and with every leaf table planned as a row scan or tiny table: equal
    distances, both equal to the oracle's."""
    rng = np.random.default_rng(40 + seed)
    costs = rand_costs(rng, 12, zeros=seed % 2 == 1)
    cost_ctx.set_costs(costs)
    kinds = (xted.GEN_RRT, xted.GEN_DEEP, xted.GEN_SHALLOW)
    trees = [gen_any(kinds[k % 3], 30 + 37 * k, 12, 932 + seed, k) for k in range(10)]
    ids = cost_ctx.pool_add(trees)
    xs, ys = ids[[0, 1, 2, 3, 4, 5, 6, 7]], ids[[9, 8, 7, 6, 5, 4, 3, 2]]
    d_on, _, _ = cost_ctx.batch(xs, ys)
    cost_ctx.force_class(xted.CLASS_AUTO, lane_enable=False)
    d_off, _, _ = cost_ctx.batch(xs, ys)
    assert np.array_equal(d_on, d_off)
    for k in range(len(xs)):
        od, _ = oracle.ted_costs(trees[int(xs[k]) - int(ids[0])], trees[int(ys[k]) - int(ids[0])],
                                 costs, want_td=False)
        assert d_on[k] == od
