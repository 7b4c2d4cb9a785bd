"""Edit mapping traced back on the host from a tree-distance table (the host
half of xted_edit_mapping): with the oracle's table, the mapping must be a
valid Tai mapping whose cost is the oracle's distance, so it is optimal."""
import numpy as np
import pytest

import oracle
import xted
from mapping_check import mapping_cost


def _pair(seed, m, n, alphabet):
    rng = np.random.default_rng(seed)
    kinds = (xted.GEN_RRT, xted.GEN_DEEP, xted.GEN_SHALLOW)
    x = xted.generate(kinds[int(rng.integers(0, 3))], m, alphabet, 7000 + seed)
    y = xted.generate(kinds[int(rng.integers(0, 3))], n, alphabet, 8000 + seed)
    return x, y


@pytest.mark.parametrize("seed", range(24))
def test_mapping_from_oracle_table_is_optimal(seed):
    rng = np.random.default_rng(seed)
    m, n = int(rng.integers(1, 80)), int(rng.integers(1, 80))
    x, y = _pair(seed, m, n, int(rng.choice([2, 4, 30])))
    dist, td, _ = oracle.ted(x, y)
    cost, pairs = xted.mapping_from_td(x, y, td)
    assert cost == dist
    assert mapping_cost(x, y, pairs) == dist


def test_mapping_tiny_trees_match_brute_force():
    for seed in range(30):
        rng = np.random.default_rng(100 + seed)
        x, y = _pair(100 + seed, int(rng.integers(1, 6)), int(rng.integers(1, 6)), 2)
        dist, td, _ = oracle.ted(x, y)
        cost, pairs = xted.mapping_from_td(x, y, td)
        assert cost == oracle.brute_force_ted(x, y) == mapping_cost(x, y, pairs)


def test_mapping_identical_trees_is_identity():
    x = xted.generate(xted.GEN_DEEP, 60, 5, 11)
    dist, td, _ = oracle.ted(x, x)
    cost, pairs = xted.mapping_from_td(x, x, td)
    assert cost == dist == 0
    assert np.array_equal(pairs, np.stack([np.arange(60), np.arange(60)], axis=1))


def test_mapping_empty_tree():
    x = xted.generate(xted.GEN_RRT, 9, 5, 12)
    empty = (np.empty(0, np.int32), np.empty(0, np.int32))
    cost, pairs = xted.mapping_from_td(x, empty, np.empty((9, 0), np.int32))
    assert cost == 9 and len(pairs) == 0


def test_mapping_check_rejects_broken_mappings():
    """The validity check is not vacuous: a crossed pair and a broken
    ancestry are caught."""
    x = (np.array([-1, 0, 0], np.int32), np.array([1, 2, 3], np.int32))  # root, two leaves
    y = (np.array([-1, 0, 1], np.int32), np.array([1, 2, 3], np.int32))  # a path
    with pytest.raises(AssertionError):
        mapping_cost(x, x, [[1, 2], [2, 1]])  # order crossed
    with pytest.raises(AssertionError):
        mapping_cost(x, y, [[1, 1], [2, 2]])  # siblings onto parent and child
    assert mapping_cost(x, y, [[0, 0], [1, 1]]) == 2
