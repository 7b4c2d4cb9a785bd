"""GPU parity: the HIP path (through the C ABI) against the oracle, bit-exact
on distances and full tree-distance tables."""
import json
import os

import numpy as np
import pytest

import oracle
import xted
from conftest import GOLDEN, golden_pairs, load_dataset_pair

pytestmark = pytest.mark.gpu


def gen(kind, n, alphabet, cfg, idx, band=(0, 0)):
    return xted.generate(kind, n, alphabet, xted.derive_seed(cfg, idx), *band)


def assert_same(ctx, x, y, threads=8):
    d, td, _ = ctx.ted(x, y)
    od, otd, _ = oracle.ted(x, y, threads=threads)
    assert d == od
    assert td.shape == otd.shape
    mism = np.argwhere(td != otd)
    assert len(mism) == 0, f"{len(mism)} TD cells differ, first {mism[:5].tolist()}"
    return d


@pytest.mark.parametrize("case", golden_pairs(), ids=lambda c: c[0])
def test_golden_reference_tables(gpu_ctx, case):
    name, x, y, td, dist, _ = case
    d, t, _ = gpu_ctx.ted(x, y)
    assert d == dist
    assert (t == td).all()


def test_rrt_100_node_pairs(gpu_ctx):
    """100-node random recursive trees, alphabet 50: distance + 100x100 table."""
    for i in range(12):
        x = gen(xted.GEN_RRT, 100, 50, 1, 2 * i)
        y = gen(xted.GEN_RRT, 100, 50, 1, 2 * i + 1)
        assert_same(gpu_ctx, x, y)


@pytest.mark.parametrize("kind", [xted.GEN_DEEP, xted.GEN_SHALLOW])
@pytest.mark.parametrize("alphabet", [200, 1000])
def test_1000_node_pairs(gpu_ctx, kind, alphabet):
    band = (60, 120) if kind == xted.GEN_DEEP else (0, 0)
    x = gen(kind, 1000, alphabet, 2, 0, band)
    y = gen(kind, 1000, alphabet, 2, 1, band)
    assert_same(gpu_ctx, x, y, threads=16)


def test_1000_node_mixed_shapes(gpu_ctx):
    x = gen(xted.GEN_DEEP, 1000, 200, 2, 10, (60, 120))
    y = gen(xted.GEN_SHALLOW, 1000, 200, 2, 11)
    assert_same(gpu_ctx, x, y, threads=16)
    assert_same(gpu_ctx, y, x, threads=16)


@pytest.mark.parametrize("cls", list(range(1, 26)) + [xted.CLASS_TINY + k for k in range(3)])
def test_every_row_class_forced(gpu_ctx, cls):
    """Every table that fits goes through one (waves, columns-per-lane) class;
    single-row tables too (lane kernel off). Results must not change a bit."""
    gpu_ctx.force_class(cls, lane_enable=False)
    try:
        for i, (kx, ky, n1, n2) in enumerate([(0, 0, 90, 70), (1, 2, 120, 150), (2, 1, 80, 130),
                                              (1, 1, 140, 110)]):
            x = gen(kx, n1, 5, 30 + cls, 2 * i)
            y = gen(ky, n2, 5, 30 + cls, 2 * i + 1)
            assert_same(gpu_ctx, x, y)
    finally:
        gpu_ctx.force_class(xted.CLASS_AUTO, lane_enable=True)


def test_lane_kernel_and_rowscan_agree(gpu_ctx):
    x = gen(xted.GEN_SHALLOW, 300, 7, 40, 0)
    y = gen(xted.GEN_RRT, 250, 7, 40, 1)
    d1, t1, _ = gpu_ctx.ted(x, y)
    gpu_ctx.force_class(xted.CLASS_AUTO, lane_enable=False)
    try:
        d2, t2, _ = gpu_ctx.ted(x, y)
    finally:
        gpu_ctx.force_class(xted.CLASS_AUTO, lane_enable=True)
    assert d1 == d2 and (t1 == t2).all()


def test_dataset_known_answers(gpu_ctx):
    with open(os.path.join(GOLDEN, "datasets", "expected.json")) as f:
        expected = json.load(f)
    for name, dist in sorted(expected.items()):
        x, y = load_dataset_pair(name)
        d, _, _ = gpu_ctx.ted(x, y, want_td=False)
        assert d == dist, name


def test_dataset_full_tables(gpu_ctx):
    for name in ["swissport_100", "python_100", "dblp_100", "swissport_1000"]:
        x, y = load_dataset_pair(name)
        assert_same(gpu_ctx, x, y)


def test_edge_cases(gpu_ctx):
    one_a = (np.array([-1], np.int32), np.array([1], np.int32))
    one_b = (np.array([-1], np.int32), np.array([2], np.int32))
    assert gpu_ctx.ted(one_a, one_a)[0] == 0
    assert gpu_ctx.ted(one_a, one_b)[0] == 1
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32))
    star = (np.array([-1, 0, 0, 0], np.int32), np.array([1, 2, 3, 2], np.int32))
    assert gpu_ctx.ted(empty, star)[0] == 4
    assert gpu_ctx.ted(star, empty)[0] == 4
    # single node against a star holding its label (root onto leaf)
    one_3 = (np.array([-1], np.int32), np.array([3], np.int32))
    assert gpu_ctx.ted(one_3, star)[0] == 3
    path = (np.arange(-1, 199, dtype=np.int32), np.arange(200, dtype=np.int32) % 5)
    wide = (np.array([-1] + [0] * 299, np.int32), np.arange(300, dtype=np.int32) % 4)
    for a in (one_a, star, path, wide):
        for b in (one_b, star, path, wide):
            assert_same(gpu_ctx, a, b)
    # identical trees: distance 0 and a zero diagonal
    x = gen(xted.GEN_RRT, 300, 9, 50, 0)
    d, td, _ = gpu_ctx.ted(x, x)
    assert d == 0 and (np.diag(td) == 0).all()


def test_structure_arrays(gpu_ctx):
    for kind in (0, 1, 2):
        for n in (1, 2, 5, 300, 2000):
            p, _ = gen(kind, n, 3, 60 + kind, n)
            orl, kr, kd = gpu_ctx.tree_structure(p)
            o = oracle.orl(p)
            assert (orl == o).all()
            k = oracle.keyroots(o)
            assert (kr == k).all()
            assert (kd == oracle.keyroot_depth(o, k)).all()


def test_invalid_inputs(gpu_ctx):
    bad = (np.array([-1, 0, 0, 1], np.int32), np.zeros(4, np.int32))
    ok = (np.array([-1, 0], np.int32), np.zeros(2, np.int32))
    with pytest.raises(xted.XtedError) as e:
        gpu_ctx.ted(bad, ok)
    assert e.value.status == xted.ERR_NOT_PREORDER
    big = (np.arange(-1, 40000, dtype=np.int32), np.zeros(40001, np.int32))
    with pytest.raises(xted.XtedError) as e:
        gpu_ctx.ted(big, ok)
    assert e.value.status == xted.ERR_TOO_LARGE
    # schedule knobs out of range are refused and leave the context as it was
    for call, arg in ((gpu_ctx.set_size_bands, 3), (gpu_ctx.set_size_bands, -1),
                      (gpu_ctx.set_light_lds, 512), (gpu_ctx.set_light_lds, 1 << 20),
                      (gpu_ctx.set_level_align, 3), (gpu_ctx.set_level_streams, 0),
                      (gpu_ctx.set_tiny_forms, -1), (gpu_ctx.set_tiny_forms, 8)):
        with pytest.raises(xted.XtedError) as e:
            call(arg)
        assert e.value.status == xted.ERR_INVALID_ARG
    # the context still works after errors
    assert gpu_ctx.ted(ok, ok)[0] == 0


def test_batch_matches_single_and_oracle(gpu_ctx):
    rng = np.random.default_rng(3)
    trees = []
    for i in range(24):
        kind = int(rng.integers(0, 3))
        trees.append(gen(kind, int(rng.integers(1, 400)), 20, 70, i))
    gpu_ctx.pool_clear()
    ids = gpu_ctx.pool_add(trees)
    xs = rng.integers(0, 24, 60).astype(np.int32)
    ys = rng.integers(0, 24, 60).astype(np.int32)
    shapes = [(len(trees[a][0]), len(trees[b][0])) for a, b in zip(xs, ys)]
    dist, tds, st = gpu_ctx.batch(ids[xs], ids[ys], td_shapes=shapes)
    assert st.tables > 0 and st.cells > 0
    for k, (a, b) in enumerate(zip(xs, ys)):
        od, otd, _ = oracle.ted(trees[a], trees[b], threads=4)
        assert dist[k] == od
        assert (tds[k] == otd).all()


def _mixed_batch(ctx, npairs, seed):
    rng = np.random.default_rng(seed)
    trees = [gen(int(rng.integers(0, 3)), int(rng.integers(2, 700)), 30, 80 + seed, i)
             for i in range(2 * npairs)]
    ctx.pool_clear()
    ids = ctx.pool_add(trees)
    shapes = [(len(trees[2 * k][0]), len(trees[2 * k + 1][0])) for k in range(npairs)]
    return trees, ids[0::2].copy(), ids[1::2].copy(), shapes


def test_batch_groups_do_not_change_results(gpu_ctx):
    """A batch cut into 1, 2, 3 or 4 independently scheduled groups (split
    from 2 pairs up): tables identical, distances in the caller's order and
    equal to the oracle's."""
    trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 37, 2)
    out = {}
    try:
        for n in (1, 2, 3, 4):
            gpu_ctx.set_batch_groups(n, 2)
            out[n] = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    finally:
        gpu_ctx.set_batch_groups(1, 256)
    for n in (2, 3, 4):
        assert np.array_equal(out[n][0], out[1][0])
        for a, b in zip(out[n][1], out[1][1]):
            assert np.array_equal(a, b)
    for k in range(0, 37, 4):
        assert out[4][0][k] == oracle.ted(trees[2 * k], trees[2 * k + 1], threads=4)[0]


def test_graph_replay_matches_direct_launches(gpu_ctx):
    """Levels captured as a graph, the graph re-launched for the same plan
    (twice) and re-captured after the pool grows: distances and tables equal
    to direct launches and to the oracle."""
    trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 24, 3)
    try:
        gpu_ctx.set_graphs(False)
        ref = gpu_ctx.batch(xs, ys, td_shapes=shapes)
        gpu_ctx.set_graphs(True)
        runs = [gpu_ctx.batch(xs, ys, td_shapes=shapes) for _ in range(3)]
        gpu_ctx.pool_add([gen(1, 900, 30, 99, k) for k in range(40)])  # pool arrays move
        runs.append(gpu_ctx.batch(xs, ys, td_shapes=shapes))
        # a different batch over the same pool, then the first one again
        other = gpu_ctx.batch(ys[::-1].copy(), xs[::-1].copy())
        runs.append(gpu_ctx.batch(xs, ys, td_shapes=shapes))
    finally:
        gpu_ctx.set_graphs(True)
    for out in runs:
        assert np.array_equal(out[0], ref[0])
        for a, b in zip(out[1], ref[1]):
            assert np.array_equal(a, b)
    for k in range(0, 24, 4):
        assert ref[0][k] == oracle.ted(trees[2 * k], trees[2 * k + 1], threads=4)[0]
        assert other[0][23 - k] == oracle.ted(trees[2 * k + 1], trees[2 * k], threads=4)[0]


def test_level_streams_do_not_change_results(gpu_ctx):
    """Launches of one level spread over 1, 2 or 4 streams: tables identical,
    distances equal to the oracle's."""
    trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 40, 1)
    out = {}
    try:
        for n in (1, 2, 4):
            gpu_ctx.set_level_streams(n)
            out[n] = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    finally:
        gpu_ctx.set_level_streams(4)
    for n in (2, 4):
        assert np.array_equal(out[n][0], out[1][0])
        for a, b in zip(out[n][1], out[1][1]):
            assert np.array_equal(a, b)
    for k in range(0, 40, 5):
        assert out[4][0][k] == oracle.ted(trees[2 * k], trees[2 * k + 1], threads=4)[0]


@pytest.mark.parametrize("rule", [(10240, 4), (20480, 16), (163840, 8)])
def test_granularity_rules_do_not_change_results(gpu_ctx, rule):
    """Autotuner candidates re-class tables; every table must stay bit-exact."""
    trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 24, 2)
    base = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    try:
        gpu_ctx.set_tuning(lds_per_wave=rule[0], max_cpl=rule[1])
        got = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(got[0], base[0])
    for a, b in zip(got[1], base[1]):
        assert np.array_equal(a, b)
    for k in range(0, 24, 6):
        od, otd, _ = oracle.ted(trees[2 * k], trees[2 * k + 1], threads=4)
        assert got[0][k] == od and np.array_equal(got[1][k], otd)


@pytest.mark.parametrize("cls", [27, 28, 29, 30, 31, 32])
def test_every_column_blocked_form_forced(gpu_ctx, cls):
    """Every multi-row table goes through one column-blocked form (blocks of
    4096 .. 128 columns, one workgroup each, chained per row through device
    memory). With 128-column blocks a 1200-column table spans 10 blocks, so
    carries and imported far values cross several block boundaries."""
    gpu_ctx.force_class(cls)
    try:
        for i, (kx, ky, n1, n2) in enumerate([(0, 1, 300, 420), (2, 0, 650, 500),
                                              (1, 2, 900, 1200), (1, 1, 700, 650)]):
            x = gen(kx, n1, 12, 50 + cls, 2 * i)
            y = gen(ky, n2, 12, 50 + cls, 2 * i + 1)
            assert_same(gpu_ctx, x, y, threads=16)
    finally:
        gpu_ctx.force_class(xted.CLASS_AUTO)


def test_column_blocked_batch_and_streams(gpu_ctx):
    """Forced column-blocked tables in a batch, launches spread over streams."""
    gpu_ctx.force_class(30)
    try:
        trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 12, 3)
        dist, tds, _ = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    finally:
        gpu_ctx.force_class(xted.CLASS_AUTO)
    for k in range(12):
        od, otd, _ = oracle.ted(trees[2 * k], trees[2 * k + 1], threads=8)
        assert dist[k] == od and np.array_equal(tds[k], otd)


def _xl_case(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_xl_golden",
                                                  os.path.join(GOLDEN, "make_xl_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(os.path.join(GOLDEN, "xl_expected.json")) as f:
        table = json.load(f)
    if name not in table:
        pytest.skip(f"no oracle checksum for {name} in xl_expected.json")
    want = table[name]
    return mod, mod.trees_of(name), want


@pytest.mark.parametrize("name", ["rrt9000_a500", "deep5000_a200", "rrt20000_a500"])
def test_large_pairs_against_oracle_checksums(gpu_ctx, name):
    """Pairs whose largest tables need the column-blocked form by width
    (RRT 9000 and 20000: more than 8192 columns) or by slot storage (deep
    5000: nesting height above 100): distance and two checksums of the full
    TD against the oracle's (tests/golden/make_xl_golden.py)."""
    mod, (x, y), want = _xl_case(name)
    d, td, st = gpu_ctx.ted(x, y)
    assert d == want["dist"]
    assert mod.td_checksums(td) == (want["td_sum"], want["td_wsum"])


def test_large_pair_transposed(gpu_ctx):
    """ted(y, x) gives the transposed table of ted(x, y) at full size."""
    _, (x, y), want = _xl_case("rrt9000_a500")
    d1, t1, _ = gpu_ctx.ted(x, y)
    d2, t2, _ = gpu_ctx.ted(y, x)
    assert d1 == d2 == want["dist"]
    assert np.array_equal(t1, t2.T)


@pytest.mark.parametrize("tail_pairs", [0, 3, 1 << 20])
def test_tail_rule_does_not_change_results(gpu_ctx, tail_pairs):
    """The latency rule for sparse top levels (off, some levels, every level)
    re-classes tables; every table must stay bit-exact."""
    trees, xs, ys, shapes = _mixed_batch(gpu_ctx, 16, 4)
    try:
        gpu_ctx.set_tail_rule(tail_pairs)
        dist, tds, _ = gpu_ctx.batch(xs, ys, td_shapes=shapes)
    finally:
        gpu_ctx.set_tail_rule(32)
    for k in range(16):
        od, otd, _ = oracle.ted(trees[2 * k], trees[2 * k + 1], threads=8)
        assert dist[k] == od and np.array_equal(tds[k], otd)
