"""All-pairs / dataset-join mode (the dimension x dimension distance matrix
the reference's test() sizes but fills only at (0, 1),
X-TED_GPU/test_TED.cu:30-54): every entry against the oracle, chunking by the
HBM bound, cost models (every ordered pair), and the command line."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import xted
import xted_text
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

SMALL = ["bolzano", "dblp_100", "python_100", "python_100_trees_2_3", "swissport_100"]


def fixture_trees(names):
    ids, trees = {}, []
    for name in names:
        base = os.path.join(GOLDEN, "datasets", name)
        trees += xted_text.load_trees(base + "_nodes.txt", base + "_adj.txt", ids)
    return trees, len(ids)


def test_dataset_matrix_against_oracle(gpu_ctx):
    trees, _ = fixture_trees(SMALL)
    assert len(trees) == 2 * len(SMALL)
    dist, st = gpu_ctx.all_pairs(trees)
    n = len(trees)
    assert st.tables > 0
    assert (np.diag(dist) == 0).all() and (dist == dist.T).all()
    for i in range(n):
        for j in range(i + 1, n):
            assert dist[i, j] == oracle.ted(trees[i], trees[j], threads=4, want_td=False)[0], (i, j)


def test_chunks_do_not_change_the_matrix(gpu_ctx):
    rng = np.random.default_rng(31)
    trees = [xted.generate(int(rng.integers(0, 3)), int(rng.integers(1, 300)), 20,
                           xted.derive_seed(970, i)) for i in range(30)]
    whole, _ = gpu_ctx.all_pairs(trees)
    chunked, st = gpu_ctx.all_pairs(trees, max_bytes=1 << 20)  # about one pair per chunk
    assert np.array_equal(whole, chunked)
    gpu_ctx.pool_clear()
    ids = gpu_ctx.pool_add(trees)
    iu, ju = np.triu_indices(30, 1)
    d, _, _ = gpu_ctx.batch(ids[iu], ids[ju])
    assert np.array_equal(d, whole[iu, ju])
    for k in range(0, len(iu), 37):
        i, j = iu[k], ju[k]
        assert whole[i, j] == oracle.ted(trees[i], trees[j], threads=4, want_td=False)[0]


def test_cost_model_fills_every_ordered_pair(gpu_ctx):
    rng = np.random.default_rng(32)
    trees = [xted.generate(int(rng.integers(0, 3)), int(rng.integers(1, 120)), 6,
                           xted.derive_seed(971, i)) for i in range(10)]
    dele = rng.integers(1, 9, 6).astype(np.int32)
    ins = rng.integers(1, 9, 6).astype(np.int32)
    ren = rng.integers(0, 12, (6, 6)).astype(np.int32)
    np.fill_diagonal(ren, 0)
    ren[2, 2] = 3  # relabelling to the same label costs something: the diagonal need not be 0
    costs = (dele, ins, ren)
    gpu_ctx.set_costs(costs)
    try:
        dist, _ = gpu_ctx.all_pairs(trees)
    finally:
        gpu_ctx.set_costs(None)
    for i in range(10):
        for j in range(10):
            assert dist[i, j] == oracle.ted_costs(trees[i], trees[j], costs, threads=4,
                                                  want_td=False)[0], (i, j)


def test_command_line_all_pairs(gpu_ctx):
    trees, _ = fixture_trees(["python_100_trees_2_3"])
    want, _ = gpu_ctx.all_pairs(trees)
    base = os.path.join(GOLDEN, "datasets", "python_100_trees_2_3")
    cli = os.path.join(ROOT, "x-ted_amd", "bin", "X-TED_GPU")
    out = subprocess.run([cli, base + "_nodes.txt", base + "_adj.txt", "--all-pairs"],
                         capture_output=True, text=True, timeout=120, check=True).stdout
    rows = [list(map(int, line.split())) for line in out.strip().split("\n")[1:]]
    assert np.array_equal(np.array(rows), want)
