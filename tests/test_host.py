"""CPU-side checks of the product library: it loads, exports every entry point
include/xted.h declares, and its host utilities (generators, validation,
text formats) behave. No compute call here touches a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import xted
import xted_text
from conftest import GOLDEN, ROOT


def declared_symbols():
    with open(os.path.join(ROOT, "include", "xted.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xted_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(xted.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    # and the binding covers exactly the declared set
    assert set(syms) == set(xted._SIGNATURES)


def test_cpp_api_symbols_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", xted.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in ["standard_ted(", "standard_ted_1(", "outermost_right_leaves(", "key_roots(",
                 "min3(", "parallel_standard_ted_test(", "node_process(", "pre_process(",
                 "splitting(", "To_Vector(", "ToInt(", "test(char*, char*)",
                 "total_milliseconds"]:
        assert name in out, name


def test_cli_usage_without_gpu():
    cli = os.path.join(ROOT, "x-ted_amd", "bin", "X-TED_GPU")
    assert os.path.exists(cli)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_status_strings():
    lib = xted.lib()
    assert lib.xted_status_string(0) == b"ok"
    assert lib.xted_status_string(2) == b"tree is not a preorder parent array"


@pytest.mark.parametrize("kind", [xted.GEN_RRT, xted.GEN_DEEP, xted.GEN_SHALLOW])
def test_generators_are_deterministic_preorder_trees(kind):
    for n in (1, 2, 17, 100, 1000):
        s = xted.derive_seed(42, n)
        p1, l1 = xted.generate(kind, n, 50, s)
        p2, l2 = xted.generate(kind, n, 50, s)
        assert (p1 == p2).all() and (l1 == l2).all()
        assert xted.validate(p1) == xted.OK
        assert len(p1) == n and l1.min() >= 0 and l1.max() < 50


def test_generator_shapes():
    depths = []
    for i in range(8):
        p, _ = xted.generate(xted.GEN_DEEP, 1000, 200, xted.derive_seed(2, i), 60, 120)
        _, depth, _ = xted.tree_stats(p)
        depths.append(depth)
        assert 60 <= depth <= 120
    for i in range(8):
        p, _ = xted.generate(xted.GEN_SHALLOW, 1000, 200, xted.derive_seed(2, 100 + i))
        _, depth, _ = xted.tree_stats(p)
        assert depth <= 8
        assert (p == 0).sum() == 40  # the root's fan-out
    # random recursive tree: parent of i uniform in [0, i) before renumbering;
    # mean depth grows like ln n
    p, _ = xted.generate(xted.GEN_RRT, 1000, 50, xted.derive_seed(1, 0))
    _, depth, _ = xted.tree_stats(p)
    assert 8 <= depth <= 30


def test_validate_rejects_non_preorder():
    assert xted.validate(np.array([-1, 0, 1, 0], np.int32)) == xted.OK
    # node 3's parent 1 is not on the path 0-2 (2's parent is 0)
    assert xted.validate(np.array([-1, 0, 0, 1], np.int32)) == xted.ERR_NOT_PREORDER
    assert xted.validate(np.array([0, 0], np.int32)) == xted.ERR_NOT_PREORDER
    assert xted.validate(np.array([-1, 1], np.int32)) == xted.ERR_NOT_PREORDER


def test_tree_stats_cells():
    # a path of n nodes has one keyroot (the root): S = n
    p = np.array([-1, 0, 1, 2], np.int32)
    assert xted.tree_stats(p) == (4, 3, 1)
    # a star: root + every leaf but the last are keyroots
    p = np.array([-1, 0, 0, 0], np.int32)
    assert xted.tree_stats(p) == (4 + 1 + 1, 1, 3)


def test_text_formats_match_reference_tokenisation():
    assert xted_text.split_labels("a b  c") == ["a", "b", ""]  # trailing token dropped
    assert xted_text.split_labels("x y ") == ["x", "y"]
    assert xted_text.parse_adjacency("[[1, 2], [], []]") == [[1, 2], [], []]
    assert xted_text.parse_adjacency("[[]]") == [[]]
    with open(os.path.join(GOLDEN, "datasets", "bolzano_nodes.txt")) as f:
        assert f.readline().endswith(" \n")


def test_load_trees_reads_every_tree_with_shared_label_ids():
    import xted_text
    ids = {}
    a = xted_text.load_trees(os.path.join(GOLDEN, "datasets", "dblp_100_nodes.txt"),
                             os.path.join(GOLDEN, "datasets", "dblp_100_adj.txt"), ids)
    b = xted_text.load_trees(os.path.join(GOLDEN, "datasets", "python_100_nodes.txt"),
                             os.path.join(GOLDEN, "datasets", "python_100_adj.txt"), ids)
    assert len(a) == 2 and len(b) == 2
    first = xted_text.load_pair(os.path.join(GOLDEN, "datasets", "dblp_100_nodes.txt"),
                                os.path.join(GOLDEN, "datasets", "dblp_100_adj.txt"))
    for (p, l), (q, m) in zip(a, first):
        assert np.array_equal(p, q) and np.array_equal(l, m)
    used = np.concatenate([t[1] for t in a + b])
    assert used.min() >= 0 and used.max() < len(ids)
