"""Python binding (ctypes) of the xted C ABI (include/xted.h).

Used by the tests and bench.py. Loads x-ted_amd/lib/libxted.so, the library
built in-tree by ``make -C x-ted_amd``; there is no fallback path: a missing
library raises at import of the first call.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libxted.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_NOT_PREORDER = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4
ERR_TOO_LARGE = 5
ERR_PARSE = 6
ERR_BUSY = 7
ERR_TIMEOUT = 8
ERR_NOT_FOUND = 9
ERR_CANCELLED = 10

CLASS_AUTO = -1
CLASS_LANE = 0
NUM_CLASSES = 36
CLASS_TINY = 33  # tiny forms 33, 34, 35: 16, 32, 64 lanes per table
TINY_SEG = (16, 32, 64)
XL_BW = (4096, 2048, 1024, 512, 256, 128)  # column-block widths of classes 27..32


def class_name(cls):
    """'leaf' (closed form for tables with a leaf keyroot) or
    'rowscan_w{W}_c{C}' (W waves, C columns per lane) or 'xl_bw{B}' (column
    blocks of B columns, one workgroup each) or 'tiny_s{S}' (tables of at
    most S = 16, 32 or 64 columns, 64 / S to a wave) or 'too_large'."""
    if cls == 0:
        return "leaf"
    if 1 <= cls <= 25:
        wi, ci = divmod(cls - 1, 5)
        return f"rowscan_w{1 << wi}_c{1 << ci}"
    if 27 <= cls < 27 + len(XL_BW):
        return f"xl_bw{XL_BW[cls - 27]}"
    if CLASS_TINY <= cls < CLASS_TINY + len(TINY_SEG):
        return f"tiny_s{TINY_SEG[cls - CLASS_TINY]}"
    return "too_large"

GEN_RRT, GEN_DEEP, GEN_SHALLOW = 0, 1, 2

_I32P = ctypes.POINTER(ctypes.c_int32)

# Every function the header declares, with (restype, argtypes).
_SIGNATURES = {
    "xted_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "xted_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "xted_last_error": (ctypes.c_char_p, []),
    "xted_status_string": (ctypes.c_char_p, [ctypes.c_int32]),
    "xted_ted": (ctypes.c_int, [ctypes.c_void_p, _I32P, _I32P, ctypes.c_int32, _I32P, _I32P,
                                ctypes.c_int32, _I32P, _I32P, ctypes.c_void_p]),
    "xted_tree_structure": (ctypes.c_int, [ctypes.c_void_p, _I32P, ctypes.c_int32, _I32P, _I32P,
                                           _I32P, _I32P]),
    "xted_edit_mapping": (ctypes.c_int, [ctypes.c_void_p, _I32P, _I32P, ctypes.c_int32, _I32P,
                                         _I32P, ctypes.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "xted_mapping_from_td": (ctypes.c_int, [_I32P, _I32P, ctypes.c_int32, _I32P, _I32P,
                                            ctypes.c_int32, _I32P, _I32P, _I32P, _I32P,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "xted_pool_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "xted_pool_clear": (ctypes.c_int, [ctypes.c_void_p]),
    "xted_pool_size": (ctypes.c_int, [ctypes.c_void_p, _I32P]),
    "xted_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P, _I32P,
                                  ctypes.c_void_p, ctypes.c_void_p]),
    "xted_batch_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P, _I32P]),
    "xted_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32]),
    "xted_force_class": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "xted_set_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_level_streams": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_batch_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "xted_set_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_level_align": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_size_bands": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_light_lds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_tiny_forms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_bracket_to_lines": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64,
                                             ctypes.c_char_p, ctypes.c_int64,
                                             ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64)]),
    "xted_lines_to_bracket": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                             ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "xted_all_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P, _I32P,
                                      ctypes.c_int64, _I32P, ctypes.c_void_p]),
    "xted_set_costs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P, _I32P]),
    "xted_set_tail_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_xl_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_set_xl_width": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "xted_class_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_double)]),
    "xted_launch_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32P, _I32P,
                                           ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), _I32P]),
    "xted_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "xted_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_uint64]),
    "xted_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_int32, _I32P]),
    "xted_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "xted_generate": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, _I32P,
                                     _I32P]),
    "xted_derive_seed": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
    "xted_validate": (ctypes.c_int, [_I32P, ctypes.c_int32]),
    "xted_tree_stats": (ctypes.c_int, [_I32P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                       _I32P, _I32P]),
}


class Stats(ctypes.Structure):
    _fields_ = [("plan_ms", ctypes.c_double), ("dp_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("cells", ctypes.c_int64),
                ("tables", ctypes.c_int64), ("levels", ctypes.c_int32),
                ("launches", ctypes.c_int32), ("xl_stale_reads", ctypes.c_int64),
                ("xl_waits", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Result(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("status", ctypes.c_int32), ("dist", ctypes.c_int32),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double)]


class XtedError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"xted status {status}: {message}")
        self.status = status


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `make -C x-ted_amd`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _p(a):
    return a.ctypes.data_as(_I32P)


def _check(status):
    if status != OK:
        raise XtedError(status, lib().xted_last_error().decode(errors="replace"))


def _tree(parent, label):
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    label = np.ascontiguousarray(label, dtype=np.int32)
    if parent.shape != label.shape or parent.ndim != 1:
        raise ValueError("parent and label must be 1-D arrays of equal length")
    return parent, label


# ---- host utilities ---------------------------------------------------------

def derive_seed(config, index):
    return int(lib().xted_derive_seed(config, index))


def mapping_from_td(x, y, td):
    """Host-only backtrace: an optimal edit mapping of x and y from their
    m x n tree-distance table td. Returns (cost, pairs)."""
    xp, xl = _tree(*x)
    yp, yl = _tree(*y)
    m, n = len(xp), len(yp)
    td = np.ascontiguousarray(td, dtype=np.int32)
    cap = max(1, min(m, n))
    mx, my = np.empty(cap, np.int32), np.empty(cap, np.int32)
    k, cost = ctypes.c_int32(), ctypes.c_int64()
    _check(lib().xted_mapping_from_td(_p(xp), _p(xl), m, _p(yp), _p(yl), n,
                                      _p(td) if m and n else None, _p(mx), _p(my),
                                      ctypes.byref(k), ctypes.byref(cost)))
    return cost.value, np.stack([mx[:k.value], my[:k.value]], axis=1)


def generate(kind, n, alphabet, seed, depth_min=0, depth_max=0):
    """Seeded preorder tree: (parent, label) int32 arrays."""
    parent = np.empty(n, np.int32)
    label = np.empty(n, np.int32)
    _check(lib().xted_generate(kind, n, alphabet, seed, depth_min, depth_max, _p(parent),
                               _p(label)))
    return parent, label


def validate(parent):
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    return lib().xted_validate(_p(parent), len(parent))


def tree_stats(parent):
    """(S = sum of keyroot subtree sizes, depth, number of keyroots)."""
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    ksum = ctypes.c_int64()
    depth = ctypes.c_int32()
    nk = ctypes.c_int32()
    _check(lib().xted_tree_stats(_p(parent), len(parent), ctypes.byref(ksum), ctypes.byref(depth),
                                 ctypes.byref(nk)))
    return ksum.value, depth.value, nk.value


# ---- device context ----------------------------------------------------------

class Context:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().xted_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().xted_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def ted(self, x, y, want_td=True):
        """x, y = (parent, label). Returns (dist, td or None, Stats)."""
        xp, xl = _tree(*x)
        yp, yl = _tree(*y)
        m, n = len(xp), len(yp)
        td = np.empty((m, n), np.int32) if (want_td and m and n) else None
        dist = ctypes.c_int32()
        st = Stats()
        _check(lib().xted_ted(self._h, _p(xp), _p(xl), m, _p(yp), _p(yl), n,
                              _p(td) if td is not None else None, ctypes.byref(dist),
                              ctypes.byref(st)))
        return dist.value, td, st

    def edit_mapping(self, x, y):
        """An optimal edit mapping of x and y: (dist, pairs) with pairs an
        (k, 2) array of matched (x node, y node), increasing in both."""
        xp, xl = _tree(*x)
        yp, yl = _tree(*y)
        m, n = len(xp), len(yp)
        cap = max(1, min(m, n))
        mx, my = np.empty(cap, np.int32), np.empty(cap, np.int32)
        k, dist = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().xted_edit_mapping(self._h, _p(xp), _p(xl), m, _p(yp), _p(yl), n, _p(mx),
                                       _p(my), ctypes.byref(k), ctypes.byref(dist)))
        return dist.value, np.stack([mx[:k.value], my[:k.value]], axis=1)

    def tree_structure(self, parent):
        """Device-computed (orl, keyroots descending, keyroot depths)."""
        parent = np.ascontiguousarray(parent, dtype=np.int32)
        n = len(parent)
        orl = np.empty(n, np.int32)
        kr = np.empty(n, np.int32)
        kd = np.empty(n, np.int32)
        K = ctypes.c_int32()
        _check(lib().xted_tree_structure(self._h, _p(parent), n, _p(orl), _p(kr), ctypes.byref(K),
                                         _p(kd)))
        return orl, kr[:K.value].copy(), kd[:K.value].copy()

    def pool_add(self, trees):
        """trees: list of (parent, label). Returns the pool ids."""
        if not trees:
            return np.empty(0, np.int32)
        sizes = np.array([len(t[0]) for t in trees], np.int32)
        parents = np.ascontiguousarray(np.concatenate([np.asarray(t[0], np.int32) for t in trees]))
        labels = np.ascontiguousarray(np.concatenate([np.asarray(t[1], np.int32) for t in trees]))
        ids = np.empty(len(trees), np.int32)
        _check(lib().xted_pool_add(self._h, len(trees), _p(sizes), _p(parents), _p(labels),
                                   _p(ids)))
        return ids

    def pool_clear(self):
        _check(lib().xted_pool_clear(self._h))

    def pool_size(self):
        n = ctypes.c_int32()
        _check(lib().xted_pool_size(self._h, ctypes.byref(n)))
        return n.value

    def batch(self, xs, ys, td_shapes=None):
        """Distances of pool pairs; td_shapes (list of (m, n), one per pair,
        None for a pair whose table is not wanted) also returns the tables."""
        xs = np.ascontiguousarray(xs, dtype=np.int32)
        ys = np.ascontiguousarray(ys, dtype=np.int32)
        dist = np.empty(len(xs), np.int32)
        st = Stats()
        tds = None
        arr = None
        if td_shapes is not None:
            if len(td_shapes) != len(xs):
                raise ValueError("td_shapes: one entry per pair")
            tds = [None if s is None else np.empty(s, np.int32) for s in td_shapes]
            arr = (ctypes.c_void_p * len(tds))(*[None if t is None else t.ctypes.data
                                                 for t in tds])
        _check(lib().xted_batch(self._h, len(xs), _p(xs), _p(ys), _p(dist),
                                ctypes.cast(arr, ctypes.c_void_p) if arr is not None else None,
                                ctypes.byref(st)))
        return dist, tds, st

    def batch_noalloc(self, xs, ys):
        """Distances of pool pairs through xted_batch_noalloc: no device
        allocation, every buffer from the library's fixed per-device array."""
        xs = np.ascontiguousarray(xs, dtype=np.int32)
        ys = np.ascontiguousarray(ys, dtype=np.int32)
        dist = np.empty(len(xs), np.int32)
        _check(lib().xted_batch_noalloc(self._h, len(xs), _p(xs), _p(ys), _p(dist)))
        return dist

    def all_pairs(self, trees, max_bytes=0):
        """The len(trees) x len(trees) distance matrix of trees ((parent,
        label) pairs) and the summed Stats; the pool is used as scratch."""
        n = len(trees)
        dist = np.zeros((n, n), np.int32)
        st = Stats()
        if n == 0:
            return dist, st
        sizes = np.array([len(t[0]) for t in trees], np.int32)
        parents = np.ascontiguousarray(np.concatenate([np.asarray(t[0], np.int32) for t in trees]))
        labels = np.ascontiguousarray(np.concatenate([np.asarray(t[1], np.int32) for t in trees]))
        _check(lib().xted_all_pairs(self._h, n, _p(sizes), _p(parents), _p(labels), max_bytes,
                                    _p(dist), ctypes.byref(st)))
        return dist, st

    def set_tuning(self, lds_per_wave=10 * 1024, max_cpl=8, min_waves=1, wide_cols=1 << 30):
        _check(lib().xted_set_tuning(self._h, lds_per_wave, max_cpl, min_waves, wide_cols))

    def set_xl_rule(self, seq_rows):
        _check(lib().xted_set_xl_rule(self._h, seq_rows))

    def set_xl_width(self, max_block_cols):
        _check(lib().xted_set_xl_width(self._h, max_block_cols))

    def set_tail_rule(self, tail_pairs):
        _check(lib().xted_set_tail_rule(self._h, tail_pairs))

    def set_level_streams(self, n):
        _check(lib().xted_set_level_streams(self._h, n))

    def set_batch_groups(self, groups, split_min=256):
        _check(lib().xted_set_batch_groups(self._h, groups, split_min))

    def set_graphs(self, on=True):
        _check(lib().xted_set_graphs(self._h, 1 if on else 0))

    def set_size_bands(self, mode=1):
        """1: tables of a launch start longest first (four size bands); 2: and
        tables that need more slot storage than a workgroup's share of LDS at
        full occupancy take launches of their own; 0 (or False): plan order."""
        _check(lib().xted_set_size_bands(self._h, int(mode)))

    def set_light_lds(self, bytes_per_wave):
        """Size-band mode 2: slot bytes per wave above which a table is heavy."""
        _check(lib().xted_set_light_lds(self._h, int(bytes_per_wave)))

    def set_tiny_forms(self, mask=1):
        """Bit k: tables of at most 16 << k columns take tiny form k (class
        CLASS_TINY + k), one per (16 << k)-lane segment of a wave."""
        _check(lib().xted_set_tiny_forms(self._h, int(mask)))

    def set_level_align(self, mode):
        """0: every pair starts at level 0; 1: every pair ends at the last
        level; 2: midway."""
        _check(lib().xted_set_level_align(self._h, mode))

    def set_costs(self, costs=None):
        """costs = (dele, ins, ren) per-label int32 arrays (alphabet, alphabet,
        alphabet x alphabet with rows = x labels), or None for unit costs."""
        if costs is None:
            _check(lib().xted_set_costs(self._h, 0, None, None, None))
            return
        dele, ins, ren = (np.ascontiguousarray(a, dtype=np.int32) for a in costs)
        a = len(dele)
        if len(ins) != a or ren.shape != (a, a):
            raise ValueError("costs: dele and ins of one length, ren square of it")
        _check(lib().xted_set_costs(self._h, a, _p(dele), _p(ins), _p(ren)))

    def set_profile(self, on=True):
        _check(lib().xted_set_profile(self._h, 1 if on else 0))

    def class_profile(self):
        """{class: (launches, tables, cells, ms)} for the classes the last call used."""
        out = {}
        for c in range(NUM_CLASSES):
            la, ta, ce = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            ms = ctypes.c_double()
            _check(lib().xted_class_profile(self._h, c, ctypes.byref(la), ctypes.byref(ta),
                                            ctypes.byref(ce), ctypes.byref(ms)))
            if la.value:
                out[c] = (la.value, ta.value, ce.value, ms.value)
        return out

    def launch_profile(self):
        """Every timed launch of the last profiled call, in launch order:
        dicts of level (-1: the leaf kernel), cls, tables, cells, start_ms
        (after the first timed launch started) and ms (device time)."""
        count = ctypes.c_int32()
        _check(lib().xted_launch_profile(self._h, -1, None, None, None, None, None, None,
                                         ctypes.byref(count)))
        out = []
        for i in range(count.value):
            lv, cl = ctypes.c_int32(), ctypes.c_int32()
            ta, ce = ctypes.c_int64(), ctypes.c_int64()
            t0, ms = ctypes.c_double(), ctypes.c_double()
            _check(lib().xted_launch_profile(self._h, i, ctypes.byref(lv), ctypes.byref(cl),
                                             ctypes.byref(ta), ctypes.byref(ce), ctypes.byref(t0),
                                             ctypes.byref(ms), None))
            out.append({"level": lv.value, "cls": cl.value, "tables": ta.value,
                        "cells": ce.value, "start_ms": t0.value, "ms": ms.value})
        return out

    def force_class(self, cls=CLASS_AUTO, lane_enable=True):
        _check(lib().xted_force_class(self._h, cls, 1 if lane_enable else 0))


def bracket_to_lines(bracket):
    """One bracket-notation tree -> (label line, adjacency line) in the
    reference's input format."""
    ln, an = ctypes.c_int64(), ctypes.c_int64()
    src = bracket.encode()
    st = lib().xted_bracket_to_lines(src, None, 0, None, 0, ctypes.byref(ln), ctypes.byref(an))
    if st not in (OK, ERR_TOO_LARGE):
        _check(st)
    lbuf, abuf = ctypes.create_string_buffer(ln.value), ctypes.create_string_buffer(an.value)
    _check(lib().xted_bracket_to_lines(src, lbuf, ln.value, abuf, an.value, ctypes.byref(ln),
                                       ctypes.byref(an)))
    return lbuf.value.decode(), abuf.value.decode()


def lines_to_bracket(labels_line, adj_line):
    """The reference's label and adjacency lines of one tree -> bracket notation."""
    need = ctypes.c_int64()
    lsrc, asrc = labels_line.encode(), adj_line.encode()
    st = lib().xted_lines_to_bracket(lsrc, asrc, None, 0, ctypes.byref(need))
    if st not in (OK, ERR_TOO_LARGE):
        _check(st)
    buf = ctypes.create_string_buffer(need.value)
    _check(lib().xted_lines_to_bracket(lsrc, asrc, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


class Server:
    """Continuous-batching server over a context's pool."""

    def __init__(self, ctx, hbm_budget, max_batch=4096):
        self._h = ctypes.c_void_p()
        self.ctx = ctx
        _check(lib().xted_server_create(ctx._h, hbm_budget, max_batch, ctypes.byref(self._h)))

    def submit(self, x_id, y_id, tag):
        _check(lib().xted_server_submit(self._h, x_id, y_id, tag))

    def poll(self, cap=1024, timeout_ms=0):
        buf = (Result * cap)()
        count = ctypes.c_int32()
        _check(lib().xted_server_poll(self._h, buf, cap, timeout_ms, ctypes.byref(count)))
        return [(r.tag, r.status, r.dist, r.latency_ms, r.queue_ms) for r in buf[:count.value]]

    def close(self):
        if self._h:
            lib().xted_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
