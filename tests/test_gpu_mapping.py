"""xted_edit_mapping end to end: the table on the device, the mapping traced
on the host; valid, and costing the oracle's distance."""
import numpy as np
import pytest

import oracle
import xted
from mapping_check import mapping_cost

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kx,ky,n,alphabet", [
    (xted.GEN_RRT, xted.GEN_RRT, 100, 50),          # 100-node random recursive trees
    (xted.GEN_DEEP, xted.GEN_SHALLOW, 1000, 200),   # 1000-node deep/narrow against shallow/wide
    (xted.GEN_SHALLOW, xted.GEN_DEEP, 700, 3),      # many matches
])
def test_edit_mapping_optimal(gpu_ctx, kx, ky, n, alphabet):
    x = xted.generate(kx, n, alphabet, 31)
    y = xted.generate(ky, n - 37, alphabet, 32)
    dist, pairs = gpu_ctx.edit_mapping(x, y)
    assert dist == oracle.ted(x, y, threads=8, want_td=False)[0]
    assert mapping_cost(x, y, pairs) == dist


def test_edit_mapping_edge_cases(gpu_ctx):
    x = xted.generate(xted.GEN_RRT, 40, 5, 33)
    d, pairs = gpu_ctx.edit_mapping(x, x)
    assert d == 0 and len(pairs) == 40
    one = (np.array([-1], np.int32), np.array([4], np.int32))
    d, pairs = gpu_ctx.edit_mapping(one, x)
    assert d == oracle.ted(one, x)[0] == mapping_cost(one, x, pairs)
    empty = (np.empty(0, np.int32), np.empty(0, np.int32))
    d, pairs = gpu_ctx.edit_mapping(empty, x)
    assert d == 40 and len(pairs) == 0
