"""Validity of an edit mapping (Tai 1979): one-to-one, increasing in the
preorder of both trees, and ancestry-preserving; its cost is renames within
the mapping plus deletions plus insertions."""
import numpy as np


def subtree_end(parent):
    """Last preorder node of each node's subtree (by a reverse sweep)."""
    n = len(parent)
    end = np.arange(n)
    for i in range(n - 1, 0, -1):
        p = parent[i]
        end[p] = max(end[p], end[i])
    return end


def mapping_cost(x, y, pairs):
    """Cost of the edit script of `pairs`, or raises AssertionError if the
    pairs are not a valid mapping."""
    (xp, xl), (yp, yl) = x, y
    m, n = len(xp), len(yp)
    pairs = np.asarray(pairs).reshape(-1, 2)
    k = len(pairs)
    if k:
        xs, ys = pairs[:, 0], pairs[:, 1]
        assert xs.min() >= 0 and xs.max() < m and ys.min() >= 0 and ys.max() < n
        assert np.all(np.diff(xs) > 0) and np.all(np.diff(ys) > 0), "not increasing in both"
        ex, ey = subtree_end(np.asarray(xp)), subtree_end(np.asarray(yp))
        for a in range(k):
            # b after a is a descendant of a in x exactly when it is in y
            in_x = xs[a + 1:] <= ex[xs[a]]
            in_y = ys[a + 1:] <= ey[ys[a]]
            assert np.array_equal(in_x, in_y), f"ancestry broken at pair {a}"
        ren = int(np.sum(np.asarray(xl)[xs] != np.asarray(yl)[ys]))
    else:
        ren = 0
    return ren + (m - k) + (n - k)
