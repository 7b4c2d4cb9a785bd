"""Pins the oracle's non-unit cost model (insert / delete / rename tables per
label; the reference hard-codes unit costs, X-TED_CPU/TED.cpp:111,115,122)
before it is trusted as the checker for the device's cost mode: unit tables
reduce to the unit oracle, an independent textbook Zhang-Shasha with costs
gives the same full tables, brute force over edit mappings and the weighted
Levenshtein closed form of path trees give the same distances, and scaling
and role-swap identities hold (a transposed operand or a swapped insert /
delete would break the last)."""
import numpy as np
import pytest

import oracle
from test_oracle import rand_tree


def rand_costs(rng, alphabet, hi=9):
    dele = rng.integers(1, hi, alphabet).astype(np.int32)
    ins = rng.integers(1, hi, alphabet).astype(np.int32)
    ren = rng.integers(0, 2 * hi, (alphabet, alphabet)).astype(np.int32)
    np.fill_diagonal(ren, 0)
    return dele, ins, ren


def unit_costs(alphabet):
    ren = np.ones((alphabet, alphabet), np.int32)
    np.fill_diagonal(ren, 0)
    return np.ones(alphabet, np.int32), np.ones(alphabet, np.int32), ren


@pytest.mark.parametrize("seed", range(6))
def test_unit_tables_match_unit_oracle(seed):
    rng = np.random.default_rng(seed)
    x, y = rand_tree(rng, 40, 6), rand_tree(rng, 35, 6, "path" if seed % 3 == 0 else "rrt")
    d0, td0, _ = oracle.ted(x, y)
    d1, td1 = oracle.ted_costs(x, y, unit_costs(6))
    assert d0 == d1 and np.array_equal(td0, td1)


@pytest.mark.parametrize("seed", range(10))
def test_costs_match_textbook_full_tables(seed):
    rng = np.random.default_rng(100 + seed)
    a = int(rng.integers(2, 8))
    x, y = rand_tree(rng, int(rng.integers(1, 60)), a), rand_tree(rng, int(rng.integers(1, 60)), a)
    c = rand_costs(rng, a)
    d0, td0 = oracle.ted_costs(x, y, c, threads=2)
    d1, td1 = oracle.textbook_costs(x, y, c)
    assert d0 == d1
    if td0 is not None:
        assert np.array_equal(td0, td1)


def test_costs_brute_force_tiny_trees():
    rng = np.random.default_rng(7)
    for _ in range(60):
        x, y = rand_tree(rng, int(rng.integers(1, 6)), 3), rand_tree(rng, int(rng.integers(1, 6)), 3)
        c = rand_costs(rng, 3, hi=5)
        assert oracle.ted_costs(x, y, c)[0] == oracle.brute_force_ted(x, y, c)


def test_costs_paths_are_weighted_levenshtein():
    rng = np.random.default_rng(8)
    for _ in range(20):
        x = rand_tree(rng, int(rng.integers(1, 30)), 5, "path")
        y = rand_tree(rng, int(rng.integers(1, 30)), 5, "path")
        c = rand_costs(rng, 5)
        lev = oracle.levenshtein(list(x[1]), list(y[1]), c)
        assert oracle.ted_costs(x, y, c)[0] == lev == oracle.textbook_costs(x, y, c)[0]


def test_costs_scale_and_swap_roles():
    rng = np.random.default_rng(9)
    for _ in range(10):
        x, y = rand_tree(rng, 30, 4), rand_tree(rng, 25, 4)
        dele, ins, ren = rand_costs(rng, 4)
        d, td = oracle.ted_costs(x, y, (dele, ins, ren))
        d3, td3 = oracle.ted_costs(x, y, (3 * dele, 3 * ins, 3 * ren))
        assert d3 == 3 * d and np.array_equal(td3, 3 * td)
        # deleting from x is inserting into y the other way round
        dt, tdt = oracle.ted_costs(y, x, (ins, dele, ren.T.copy()))
        assert dt == d and np.array_equal(tdt, td.T)


def test_costs_empty_trees_sum_the_other_side():
    rng = np.random.default_rng(10)
    x = rand_tree(rng, 12, 4)
    dele, ins, ren = rand_costs(rng, 4)
    empty = (np.empty(0, np.int32), np.empty(0, np.int32))
    assert oracle.ted_costs(x, empty, (dele, ins, ren))[0] == int(dele[x[1]].sum())
    assert oracle.ted_costs(empty, x, (dele, ins, ren))[0] == int(ins[x[1]].sum())


def _subtree_min_sweep(parent, orl, w):
    """Subtree minima m[c] = min(w[k] for k in [c, orl[c]]) by one right-to-left
    sweep with a stack of pending sibling values: R[c] = min over c's subtree
    and its later siblings' subtrees; a node pops its next sibling's R and
    pushes its own R when it has a previous sibling. Returns (m, deepest
    stack)."""
    n = len(parent)
    m = [0] * n
    stack, deepest, r_prev = [], 0, 0
    for c in range(n - 1, -1, -1):
        internal = orl[c] > c
        has_prev = c > 0 and parent[c] != c - 1
        has_next = c > 0 and orl[c] < orl[parent[c]]
        mc = min(w[c], r_prev) if internal else w[c]
        rc = min(mc, stack.pop()) if has_next else mc
        if has_prev:
            stack.append(rc)
            deepest = max(deepest, len(stack))
        r_prev = rc
        m[c] = mc
    assert not stack
    return m, deepest


@pytest.mark.parametrize("seed", range(12))
def test_cost_leaf_lines_closed_form(seed):
    """Tables with a one-node side under a cost model: for a leaf v of x and
    any node c of y, TD[v][c] = sum ins(sub_y c) + min(del(v), min over k in
    sub_y(c) of ren(v, k) - ins(k)) (v deleted, or mapped onto one node k of
    the subtree, the rest inserted), and symmetrically for a leaf of y. Every
    leaf row and column of the oracle's full table must follow it; the
    subtree minima also come out of the single right-to-left sweep the device
    uses, checked here against their definition."""
    rng = np.random.default_rng(100 + seed)
    a = int(rng.integers(2, 8))
    kinds = ("rrt", "path", "star")
    x = rand_tree(rng, int(rng.integers(1, 45)), a, kinds[seed % 3])
    y = rand_tree(rng, int(rng.integers(1, 45)), a, kinds[(seed // 3) % 3])
    dele, ins, ren = rand_costs(rng, a)
    if seed % 4 == 1:
        dele[:] = 0  # free deletions
    _, td = oracle.ted_costs(x, y, (dele, ins, ren))
    for (rp, rl), (tp, tl), line_of_x in ((x, y, True), (y, x, False)):
        rp, rl, tp, tl = (np.asarray(v) for v in (rp, rl, tp, tl))
        if len(rp) == 0 or len(tp) == 0:
            continue
        ro, to = np.asarray(oracle.orl(rp)), np.asarray(oracle.orl(tp))
        # costs of the other tree's nodes when v is mapped or not
        keep = ins if line_of_x else dele
        for v in np.flatnonzero(ro == np.arange(len(ro))):
            w = [int((ren[rl[v], tl[k]] if line_of_x else ren[tl[k], rl[v]]) - keep[tl[k]])
                 for k in range(len(tp))]
            m, _ = _subtree_min_sweep(tp, to, w)
            dv = int(dele[rl[v]] if line_of_x else ins[rl[v]])
            for c in range(len(tp)):
                assert m[c] == min(w[c:to[c] + 1])
                want = int(keep[tl[c:to[c] + 1]].sum()) + min(dv, m[c])
                got = td[v, c] if line_of_x else td[c, v]
                assert got == want, (seed, line_of_x, v, c)
