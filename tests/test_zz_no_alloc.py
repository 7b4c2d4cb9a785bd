"""xted_batch_noalloc (no device allocation; buffers carved from the
library's fixed per-device array) against xted_batch on the same pool pairs:
two calls in a row, the second with fewer and smaller pairs, one call under
a forced column-blocked form (its cross-workgroup words come from the array
too), and the refusal of pairs that do not fit the array."""
import numpy as np
import pytest

import oracle
import xted

pytestmark = pytest.mark.gpu


def pool_trees():
    trees = []
    for k in range(24):
        kind = (xted.GEN_RRT, xted.GEN_DEEP, xted.GEN_SHALLOW)[k % 3]
        n = 60 + 17 * k
        trees.append(xted.generate(kind, n, 30, xted.derive_seed(71, k)))
    return trees


def test_batch_noalloc_matches_batch():
    trees = pool_trees()
    rng = np.random.default_rng(5)
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(trees)
        xa, ya = ids[rng.integers(0, 24, 40)], ids[rng.integers(0, 24, 40)]
        small = ids[:8]
        xb, yb = small[rng.integers(0, 8, 6)], small[rng.integers(0, 8, 6)]
        print("noalloc: 40 pairs", flush=True)
        ref_a, _, _ = ctx.batch(xa, ya)
        got_a = ctx.batch_noalloc(xa, ya)
        assert np.array_equal(got_a, ref_a)
        print("noalloc: 6 smaller pairs", flush=True)
        ref_b, _, _ = ctx.batch(xb, yb)
        got_b = ctx.batch_noalloc(xb, yb)
        assert np.array_equal(got_b, ref_b)
        for k in range(len(xb)):
            od, _, _ = oracle.ted(trees[int(xb[k])], trees[int(yb[k])], want_td=False)
            assert got_b[k] == od
        # A forced column-blocked form (1024-column blocks): its words are
        # carved from the array and cleared there.
        print("noalloc: forced column-blocked form", flush=True)
        ctx.force_class(29)
        try:
            ref_c, _, _ = ctx.batch(xa[:12], ya[:12])
            got_c = ctx.batch_noalloc(xa[:12], ya[:12])
        finally:
            ctx.force_class(-1)
        assert np.array_equal(got_c, ref_c) and np.array_equal(got_c, ref_a[:12])


def test_batch_noalloc_refuses_what_does_not_fit():
    # Two 3000-node trees: 72 MB of tables alone, above the 64 MiB array.
    big = [xted.generate(xted.GEN_RRT, 3000, 50, xted.derive_seed(72, k)) for k in range(2)]
    with xted.Context(0) as ctx:
        ids = ctx.pool_add(big)
        with pytest.raises(xted.XtedError) as e:
            ctx.batch_noalloc(ids[:1], ids[1:])
        assert e.value.status == xted.ERR_TOO_LARGE
        assert "scratch" in str(e.value)
        # The context stays usable, and so does the array.
        small = ctx.pool_add([xted.generate(xted.GEN_RRT, 100, 50, xted.derive_seed(72, 9))] * 2)
        d = ctx.batch_noalloc(small[:1], small[1:])
        assert d[0] == 0
