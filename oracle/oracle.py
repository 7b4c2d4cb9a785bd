"""Test infrastructure: ctypes binding of the CPU oracle (lib/liboracle.so)
and, when it was built, of the reference's own X-TED_CPU (_ref/libxted_ref.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module. Also holds two pure-Python checks for tiny inputs: brute force over
edit mappings and Levenshtein distance (the closed forms of path / star trees).
"""
import ctypes
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libxted_ref.so")
REF_CLI = os.path.join(HERE, "_ref", "X-TED_CPU")

_I32P = ctypes.POINTER(ctypes.c_int32)
_lib = None
_ref = None


class _Costs(ctypes.Structure):
    _fields_ = [("alphabet", ctypes.c_int32), ("dele", _I32P), ("ins", _I32P), ("ren", _I32P)]


def _costs(costs):
    """costs = (dele[A], ins[A], ren[A, A]) int32 arrays, or None for unit
    costs. Returns (struct pointer or None, arrays kept alive)."""
    if costs is None:
        return None, ()
    dele, ins, ren = (np.ascontiguousarray(a, dtype=np.int32) for a in costs)
    a = len(dele)
    assert len(ins) == a and ren.shape == (a, a)
    c = _Costs(a, dele.ctypes.data_as(_I32P), ins.ctypes.data_as(_I32P), ren.ctypes.data_as(_I32P))
    return ctypes.pointer(c), (dele, ins, ren, c)


def _p(a):
    return a.ctypes.data_as(_I32P)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing; build with `make -C oracle`")
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_ted.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P, ctypes.c_int32,
                                    ctypes.c_int32, _I32P, ctypes.POINTER(ctypes.c_double)]
        _lib.zs_textbook.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P, ctypes.c_int32,
                                     _I32P]
        _lib.oracle_ted_costs.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P,
                                          ctypes.c_int32, ctypes.POINTER(_Costs), ctypes.c_int32,
                                          _I32P, ctypes.POINTER(ctypes.c_double)]
        _lib.zs_textbook_costs.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P,
                                           ctypes.c_int32, ctypes.POINTER(_Costs), _I32P]
        _lib.oracle_orl.argtypes = [_I32P, ctypes.c_int32, _I32P]
        _lib.oracle_keyroots.argtypes = [_I32P, ctypes.c_int32, _I32P, _I32P]
        _lib.oracle_keyroot_depth.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P]
        _lib.oracle_levels.argtypes = [_I32P, ctypes.c_int32, _I32P, ctypes.c_int32, _I32P, _I32P,
                                       _I32P]
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_LIB} missing; build with `make -C oracle` where the "
                               "reference sources are present")
        _ref = ctypes.CDLL(REF_LIB)
        _ref.ref_ted.argtypes = [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P, ctypes.c_int32,
                                 ctypes.c_int32, _I32P, ctypes.POINTER(ctypes.c_double)]
        _ref.ref_orl.argtypes = [_I32P, ctypes.c_int32, _I32P]
        _ref.ref_key_roots.argtypes = [_I32P, ctypes.c_int32, _I32P, _I32P]
    return _ref


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def ted(x, y, threads=1, want_td=True):
    """Oracle restatement of X-TED_CPU: (dist, td, compute_ms)."""
    xp, xl = map(_arr, x)
    yp, yl = map(_arr, y)
    m, n = len(xp), len(yp)
    td = np.empty((m, n), np.int32) if (want_td and m and n) else None
    ms = ctypes.c_double()
    d = _load().oracle_ted(_p(xp), _p(xl), m, _p(yp), _p(yl), n, threads,
                           _p(td) if td is not None else None, ctypes.byref(ms))
    if d < 0:
        raise ValueError("oracle rejected the input")
    return d, td, ms.value


def ted_costs(x, y, costs, threads=1, want_td=True):
    """oracle_ted under a cost model (dele, ins, ren arrays; None = unit):
    (dist, td)."""
    xp, xl = map(_arr, x)
    yp, yl = map(_arr, y)
    m, n = len(xp), len(yp)
    td = np.empty((m, n), np.int32) if (want_td and m and n) else None
    cp, keep = _costs(costs)
    d = _load().oracle_ted_costs(_p(xp), _p(xl), m, _p(yp), _p(yl), n, cp, threads,
                                 _p(td) if td is not None else None, None)
    if d < 0:
        raise ValueError("oracle rejected the input")
    return d, td


def textbook_costs(x, y, costs):
    xp, xl = map(_arr, x)
    yp, yl = map(_arr, y)
    m, n = len(xp), len(yp)
    td = np.empty((m, n), np.int32) if (m and n) else None
    cp, keep = _costs(costs)
    d = _load().zs_textbook_costs(_p(xp), _p(xl), m, _p(yp), _p(yl), n, cp,
                                  _p(td) if td is not None else None)
    return d, td


def textbook(x, y):
    xp, xl = map(_arr, x)
    yp, yl = map(_arr, y)
    m, n = len(xp), len(yp)
    td = np.empty((m, n), np.int32) if (m and n) else None
    d = _load().zs_textbook(_p(xp), _p(xl), m, _p(yp), _p(yl), n,
                            _p(td) if td is not None else None)
    return d, td


def ref_ted(x, y, threads=1, want_td=True):
    """The compiled reference X-TED_CPU: (dist, td, compute_ms)."""
    xp, xl = map(_arr, x)
    yp, yl = map(_arr, y)
    m, n = len(xp), len(yp)
    td = np.empty((m, n), np.int32) if (want_td and m and n) else None
    ms = ctypes.c_double()
    d = _load_ref().ref_ted(_p(xp), _p(xl), m, _p(yp), _p(yl), n, threads,
                            _p(td) if td is not None else None, ctypes.byref(ms))
    return d, td, ms.value


def orl(parent):
    p = _arr(parent)
    out = np.empty(len(p), np.int32)
    if _load().oracle_orl(_p(p), len(p), _p(out)) != 0:
        raise ValueError("invalid parent array")
    return out


def keyroots(orl_arr):
    o = _arr(orl_arr)
    kr = np.empty(len(o), np.int32)
    K = ctypes.c_int32()
    _load().oracle_keyroots(_p(o), len(o), _p(kr), ctypes.byref(K))
    return kr[:K.value].copy()


def keyroot_depth(orl_arr, kr):
    o, k = _arr(orl_arr), _arr(kr)
    d = np.empty(len(k), np.int32)
    _load().oracle_keyroot_depth(_p(o), _p(k), len(k), _p(d))
    return d


def levels(xp, yp):
    xp, yp = _arr(xp), _arr(yp)
    K, L = ctypes.c_int32(), ctypes.c_int32()
    out = np.empty(len(xp) * len(yp), np.int32)
    _load().oracle_levels(_p(xp), len(xp), _p(yp), len(yp), _p(out), ctypes.byref(K),
                          ctypes.byref(L))
    return out[:K.value * L.value].reshape(K.value, L.value)


def ref_orl(parent):
    p = _arr(parent)
    out = np.empty(len(p), np.int32)
    _load_ref().ref_orl(_p(p), len(p), _p(out))
    return out


def ref_key_roots(orl_arr):
    o = _arr(orl_arr)
    kr = np.empty(len(o), np.int32)
    K = ctypes.c_int32()
    _load_ref().ref_key_roots(_p(o), len(o), _p(kr), ctypes.byref(K))
    return kr[:K.value].copy()


# ---- pure-Python checks for tiny inputs -----------------------------------------

def levenshtein(a, b, costs=None):
    """Edit distance of two sequences; with costs = (dele, ins, ren) the
    weighted one (deleting a costs dele[a], inserting b ins[b], replacing a
    by b ren[a][b])."""
    dl = (lambda c: 1) if costs is None else (lambda c: int(costs[0][c]))
    il = (lambda c: 1) if costs is None else (lambda c: int(costs[1][c]))
    rl = (lambda p, q: int(p != q)) if costs is None else (lambda p, q: int(costs[2][p][q]))
    prev = [0]
    for cb in b:
        prev.append(prev[-1] + il(cb))
    for ca in a:
        cur = [prev[0] + dl(ca)]
        for j, cb in enumerate(b, 1):
            cur.append(min(prev[j] + dl(ca), cur[j - 1] + il(cb), prev[j - 1] + rl(ca, cb)))
        prev = cur
    return prev[-1]


def _ancestors(parent):
    n = len(parent)
    anc = [set() for _ in range(n)]
    for i in range(1, n):
        anc[i] = anc[parent[i]] | {parent[i]}
    return anc


def brute_force_ted(x, y, costs=None):
    """Minimum over all valid edit mappings (Tai 1979): a one-to-one partial
    map M that preserves preorder order and ancestry; cost = renames in M +
    deletions + insertions (unit, or the (dele, ins, ren) cost tables).
    Exponential; use for trees of <= 6 nodes."""
    xp, xl = [list(map(int, a)) for a in x]
    yp, yl = [list(map(int, a)) for a in y]
    m, n = len(xp), len(yp)
    dl = [1] * m if costs is None else [int(costs[0][a]) for a in xl]
    il = [1] * n if costs is None else [int(costs[1][b]) for b in yl]
    ren = (lambda i, j: int(xl[i] != yl[j])) if costs is None else \
        (lambda i, j: int(costs[2][xl[i]][yl[j]]))
    ax, ay = _ancestors(xp), _ancestors(yp)
    best = sum(dl) + sum(il)
    for k in range(0, min(m, n) + 1):
        for xs in itertools.combinations(range(m), k):
            for ys in itertools.combinations(range(n), k):
                ok = True
                for a in range(k):
                    for b in range(a + 1, k):
                        if (xs[a] in ax[xs[b]]) != (ys[a] in ay[ys[b]]):
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    continue
                cost = (sum(ren(i, j) for i, j in zip(xs, ys)) + sum(dl) - sum(dl[i] for i in xs)
                        + sum(il) - sum(il[j] for j in ys))
                best = min(best, cost)
    return best
