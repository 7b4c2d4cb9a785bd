"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something that does not share its
code: the reference's own X-TED_CPU compiled from its sources (golden
fixtures written by it, and the live build where present), the dataset known
answers, an independent textbook Zhang-Shasha, brute force over edit
mappings, Levenshtein closed forms and metric identities. A planted-mistake
test shows the checks catch a plausible error.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_pairs, load_dataset_pair


def rand_tree(rng, n, alphabet, kind="rrt"):
    """Independent numpy generator (not the product's) for oracle checks."""
    parent0 = [-1] + [int(rng.integers(0, i)) for i in range(1, n)]
    if kind == "path":
        parent0 = [-1] + list(range(n - 1))
    elif kind == "star":
        parent0 = [-1] + [0] * (n - 1)
    ch = [[] for _ in range(n)]
    for i in range(1, n):
        ch[parent0[i]].append(i)
    order, st = [], [0]
    while st:
        v = st.pop()
        order.append(v)
        st.extend(reversed(ch[v]))
    nid = {v: i for i, v in enumerate(order)}
    p = np.full(n, -1, np.int32)
    for v in range(1, n):
        p[nid[v]] = nid[parent0[v]]
    return p, rng.integers(0, alphabet, n).astype(np.int32)


def subtree_sizes(parent):
    n = len(parent)
    size = np.ones(n, np.int64)
    for i in range(n - 1, 0, -1):
        size[parent[i]] += size[i]
    return size


# ---- structural functions -------------------------------------------------------

def test_orl_closed_form_and_reference():
    rng = np.random.default_rng(7)
    for _ in range(50):
        p, _ = rand_tree(rng, int(rng.integers(1, 80)), 5)
        o = oracle.orl(p)
        assert (o == np.arange(len(p)) + subtree_sizes(p) - 1).all()
        if oracle.ref_available():
            assert (o == oracle.ref_orl(p)).all()


def test_keyroots_closed_form_and_reference():
    rng = np.random.default_rng(8)
    for _ in range(50):
        p, _ = rand_tree(rng, int(rng.integers(1, 80)), 5)
        o = oracle.orl(p)
        kr = oracle.keyroots(o)
        n = len(p)
        last_child = {}
        for i in range(1, n):
            last_child[int(p[i])] = i
        expect = sorted([0] + [i for i in range(1, n) if last_child[int(p[i])] != i], reverse=True)
        assert list(kr) == expect
        if oracle.ref_available():
            assert list(kr) == list(oracle.ref_key_roots(o))


def nesting_height(parent, kr):
    """Height of each keyroot in the keyroot-nesting tree, by recursion."""
    n = len(parent)
    o = np.arange(n) + subtree_sizes(parent) - 1
    krs = set(int(k) for k in kr)

    def height(k):
        inner = [t for t in krs if k < t <= o[k]]
        return 0 if not inner else 1 + max(height(t) for t in inner)
    return [height(int(k)) for k in kr]


def test_keyroot_depth_matches_nesting_height():
    rng = np.random.default_rng(9)
    for _ in range(30):
        p, _ = rand_tree(rng, int(rng.integers(1, 60)), 5)
        o = oracle.orl(p)
        kr = oracle.keyroots(o)
        assert list(oracle.keyroot_depth(o, kr)) == nesting_height(p, kr)


def test_levels_are_a_topological_order():
    """Every forest cell's TD[i][j] is written by table (kr(i), kr(j)) at a
    strictly lower level than the table that reads it."""
    rng = np.random.default_rng(10)
    for _ in range(12):
        xp, _ = rand_tree(rng, int(rng.integers(2, 40)), 3)
        yp, _ = rand_tree(rng, int(rng.integers(2, 40)), 3)
        lev = oracle.levels(xp, yp)
        xo, yo = oracle.orl(xp), oracle.orl(yp)
        xk, yk = list(oracle.keyroots(xo)), list(oracle.keyroots(yo))
        top_x = {int(xo[k]): a for a, k in enumerate(xk)}  # path leaf -> keyroot index
        top_y = {int(yo[k]): b for b, k in enumerate(yk)}
        for a, i0 in enumerate(xk):
            for b, j0 in enumerate(yk):
                for i in range(i0, xo[i0] + 1):
                    for j in range(j0, yo[j0] + 1):
                        if xo[i] == xo[i0] and yo[j] == yo[j0]:
                            continue
                        assert lev[top_x[int(xo[i])], top_y[int(yo[j])]] < lev[a, b]


# ---- distances and tables ---------------------------------------------------------

@pytest.mark.parametrize("case", golden_pairs(), ids=lambda c: c[0])
def test_oracle_matches_reference_golden(case):
    name, x, y, td, dist, _ = case
    d, t, _ = oracle.ted(x, y, threads=2)
    assert d == dist
    assert (t == td).all()


def test_golden_generator_is_reproducible():
    import xted
    for name, x, y, _, _, gen in golden_pairs():
        for tree, g in zip((x, y), gen):
            kind, n, alphabet, seed = (int(v) for v in g)
            p, lab = xted.generate(kind, n, alphabet, seed)
            assert (p == tree[0]).all() and (lab == tree[1]).all(), name


def test_dataset_known_answers():
    with open(os.path.join(GOLDEN, "datasets", "expected.json")) as f:
        expected = json.load(f)
    for name, dist in sorted(expected.items()):
        x, y = load_dataset_pair(name)
        if len(x[0]) * len(y[0]) > 1200 * 1200:
            continue  # the 2000-node pair runs in the GPU suite
        d, _, _ = oracle.ted(x, y, threads=4, want_td=False)
        assert d == dist, name


def test_textbook_matches_oracle_full_tables():
    rng = np.random.default_rng(11)
    for _ in range(40):
        x = rand_tree(rng, int(rng.integers(1, 50)), 4)
        y = rand_tree(rng, int(rng.integers(1, 50)), 4)
        d1, t1, _ = oracle.ted(x, y)
        d2, t2 = oracle.textbook(x, y)
        assert d1 == d2 and (t1 == t2).all()


def test_live_reference_matches_oracle():
    if not oracle.ref_available():
        pytest.skip("reference build not present")
    rng = np.random.default_rng(12)
    for _ in range(20):
        x = rand_tree(rng, int(rng.integers(1, 70)), 6)
        y = rand_tree(rng, int(rng.integers(1, 70)), 6)
        d1, t1, _ = oracle.ted(x, y, threads=3)
        d2, t2, _ = oracle.ref_ted(x, y, threads=3)
        assert d1 == d2 and (t1 == t2).all()


def test_brute_force_tiny_trees():
    rng = np.random.default_rng(13)
    for _ in range(120):
        x = rand_tree(rng, int(rng.integers(1, 6)), 3)
        y = rand_tree(rng, int(rng.integers(1, 6)), 3)
        bf = oracle.brute_force_ted(x, y)
        assert oracle.ted(x, y)[0] == bf
        assert oracle.textbook(x, y)[0] == bf


def test_subtree_tables_are_subtree_distances():
    """TD[i][j] = TED(subtree_x(i), subtree_y(j)) by brute force on tiny trees."""
    rng = np.random.default_rng(14)

    def subtree(tree, i):
        p, lab = tree
        size = subtree_sizes(p)[i]
        idx = np.arange(i, i + size)
        sp = np.array([-1] + [int(p[k]) - i for k in idx[1:]], np.int32)
        return sp, lab[idx]
    for _ in range(25):
        x = rand_tree(rng, int(rng.integers(2, 7)), 3)
        y = rand_tree(rng, int(rng.integers(2, 7)), 3)
        _, td, _ = oracle.ted(x, y)
        for i in range(len(x[0])):
            for j in range(len(y[0])):
                assert td[i, j] == oracle.brute_force_ted(subtree(x, i), subtree(y, j))


def test_paths_reduce_to_levenshtein():
    rng = np.random.default_rng(15)
    for _ in range(30):
        x = rand_tree(rng, int(rng.integers(1, 40)), 4, kind="path")
        y = rand_tree(rng, int(rng.integers(1, 40)), 4, kind="path")
        assert oracle.ted(x, y)[0] == oracle.levenshtein(list(x[1]), list(y[1]))


def star_ted(x, y):
    """Closed form for two stars (a root with leaf children). A mapping either
    maps root to root (rename + Levenshtein of the child sequences), maps one
    root onto a leaf of the other star (every other node is deleted or
    inserted), or maps nothing."""
    m, n = len(x[1]), len(y[1])
    rx, cx, ry, cy = int(x[1][0]), list(x[1][1:]), int(y[1][0]), list(y[1][1:])
    best = m + n
    best = min(best, int(rx != ry) + oracle.levenshtein(cx, cy))
    for leaf in cy:
        best = min(best, int(rx != leaf) + (m - 1) + (n - 1))
    for leaf in cx:
        best = min(best, int(leaf != ry) + (m - 1) + (n - 1))
    return best


def test_stars_closed_form():
    rng = np.random.default_rng(16)
    for _ in range(60):
        x = rand_tree(rng, int(rng.integers(1, 40)), 4, kind="star")
        y = rand_tree(rng, int(rng.integers(1, 40)), 4, kind="star")
        assert oracle.ted(x, y)[0] == star_ted(x, y)
    # the root-onto-leaf case: a single node against a star holding its label
    x = (np.array([-1], np.int32), np.array([3], np.int32))
    y = (np.array([-1, 0, 0, 0], np.int32), np.array([1, 2, 3, 2], np.int32))
    assert oracle.ted(x, y)[0] == star_ted(x, y) == 3 == oracle.brute_force_ted(x, y)


def test_metric_properties():
    rng = np.random.default_rng(17)
    trees = [rand_tree(rng, int(rng.integers(1, 30)), 3) for _ in range(8)]
    D = np.zeros((8, 8), np.int64)
    for a in range(8):
        for b in range(8):
            d, td, _ = oracle.ted(trees[a], trees[b])
            D[a, b] = d
            m, n = len(trees[a][0]), len(trees[b][0])
            assert abs(m - n) <= d <= m + n - 1
            if a == b:
                assert d == 0 and (np.diag(td) == 0).all()
            _, tdt, _ = oracle.ted(trees[b], trees[a])
            assert (td == tdt.T).all()
    for a in range(8):
        for b in range(8):
            for c in range(8):
                assert D[a, c] <= D[a, b] + D[b, c]


def test_empty_trees():
    x = (np.zeros(0, np.int32), np.zeros(0, np.int32))
    y = (np.array([-1, 0, 0], np.int32), np.array([1, 2, 3], np.int32))
    assert oracle.ted(x, y)[0] == 3
    assert oracle.ted(y, x)[0] == 3


def test_planted_mistake_is_caught():
    """A dropped rename cost (a plausible bug) must fail the brute-force and
    Levenshtein checks above: shown here on the same inputs."""
    rng = np.random.default_rng(18)
    caught = False
    for _ in range(50):
        x = rand_tree(rng, int(rng.integers(2, 6)), 2)
        y = rand_tree(rng, int(rng.integers(2, 6)), 2)
        d, _, _ = oracle.ted(x, y)
        # a buggy variant that ignores labels equals the distance of unlabeled trees
        zx = (x[0], np.zeros_like(x[1]))
        zy = (y[0], np.zeros_like(y[1]))
        buggy = oracle.ted(zx, zy)[0]
        if buggy != oracle.brute_force_ted(x, y):
            caught = True
        assert d == oracle.brute_force_ted(x, y)
    assert caught


def test_leaf_rows_closed_form():
    """TED({v}, T) = |T| - [label(v) occurs in T]: every leaf row and leaf
    column of the oracle's TD (the cells the leaf kernel writes in closed
    form) must follow it."""
    import xted
    rng = np.random.default_rng(11)
    for _ in range(20):
        x = xted.generate(int(rng.integers(0, 3)), int(rng.integers(1, 70)), 4,
                          int(rng.integers(1 << 30)))
        y = xted.generate(int(rng.integers(0, 3)), int(rng.integers(1, 70)), 4,
                          int(rng.integers(1 << 30)))
        _, td, _ = oracle.ted(x, y, threads=2)
        ox, oy = np.asarray(oracle.orl(x[0])), np.asarray(oracle.orl(y[0]))
        for i in np.flatnonzero(ox == np.arange(len(ox))):
            for c in range(len(oy)):
                want = oy[c] - c + 1 - int(x[1][i] in y[1][c:oy[c] + 1])
                assert td[i, c] == want
        for j in np.flatnonzero(oy == np.arange(len(oy))):
            for r in range(len(ox)):
                want = ox[r] - r + 1 - int(y[1][j] in x[1][r:ox[r] + 1])
                assert td[r, j] == want
