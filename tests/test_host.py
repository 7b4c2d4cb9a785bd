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


def _synthetic_files():
    d = os.path.join(GOLDEN, "datasets")
    with open(os.path.join(d, "xl_synthetic_1000_tree.txt")) as f:
        brackets = [l for l in f.read().split("\n") if l.strip()]
    with open(os.path.join(d, "xl_synthetic_1000_nodes.txt")) as f:
        nodes = f.read().split("\n")
    with open(os.path.join(d, "xl_synthetic_1000_adj.txt")) as f:
        adjs = f.read().split("\n")
    return brackets, nodes, adjs


def test_bracket_notation_matches_the_reference_files():
    """The reference ships each synthetic dataset three ways: bracket
    notation, label lines and adjacency lines (the same two 1000-node trees);
    converting its bracket file gives its other two files byte for byte, and
    converting those back gives the bracket lines."""
    brackets, nodes, adjs = _synthetic_files()
    assert len(brackets) == 2
    for k, b in enumerate(brackets):
        labels_line, adj_line = xted.bracket_to_lines(b)
        assert labels_line == nodes[k]
        assert adj_line == adjs[k]
        assert xted.lines_to_bracket(nodes[k], adjs[k]) == b


def test_bracket_round_trips_and_rejects_malformed_text():
    for b in ["{a}", "{a{b}{c{d}}}", "{x{x{x}}{y}}", "{root{c1{g1}{g2}}{c2}}"]:
        assert xted.lines_to_bracket(*xted.bracket_to_lines(b)) == b
    assert xted.bracket_to_lines("{a{b}{c}}") == ("a b c ", "[[1, 2], [], []]")
    for bad in ["", "{a", "a}", "{a}{b}", "{a}}", "{a{b}} x"]:
        with pytest.raises(xted.XtedError):
            xted.bracket_to_lines(bad)


def test_command_line_bracket_conversion(tmp_path):
    brackets, nodes, adjs = _synthetic_files()
    cli = os.path.join(ROOT, "x-ted_amd", "bin", "X-TED_GPU")
    out_n, out_a = tmp_path / "n.txt", tmp_path / "a.txt"
    subprocess.run([cli, "--bracket-to-lines",
                    os.path.join(GOLDEN, "datasets", "xl_synthetic_1000_tree.txt"),
                    str(out_n), str(out_a)], check=True, timeout=60)
    assert out_n.read_text() == "\n".join(nodes)
    assert out_a.read_text() == "\n".join(adjs)


def test_autotune_accepts_only_non_overlapping_wins(monkeypatch):
    """The tuner's acceptance rule on scripted step times (no device): a
    candidate whose steps overlap the best rule's fastest step is logged but
    not taken; one whose every step is faster becomes the best, and later
    candidates are compared with it."""
    import autotune
    times = {8: [1.00, 1.08, 1.12, 1.02, 1.05, 1.10],  # the default rule (6 steps)
             16: [0.95, 0.97, 1.01],                     # overlaps 1.00: rejected
             4: [0.90, 0.92, 0.93],                      # every step < 1.00: new best
             2: [0.89, 0.91, 0.95]}                      # overlaps 0.90: rejected
    dist = np.arange(5, dtype=np.int32)

    def fake_steps(ctx, xs, ys, rule, reps):
        return list(times[rule["max_cpl"]]), dist

    monkeypatch.setattr(autotune, "steps", fake_steps)
    monkeypatch.setattr(autotune, "apply", lambda ctx, rule: None)
    monkeypatch.setattr(autotune, "FINE_SPAN", 0)
    logged = []
    best, results, spread = autotune.tune(None, None, None, stages=[{"max_cpl": [16, 4, 2]}],
                                          log=logged.append, confirm_top=1)
    assert best["max_cpl"] == 4
    assert [r["max_cpl"] for r, _ in results] == [8, 16, 4, 2]
    assert spread == pytest.approx(0.12)
    assert sum("new best" in m for m in logged) == 1


def test_autotune_confirmation_takes_lowest_median(monkeypatch):
    """After the search, the best rule and the next fastest candidates are
    re-timed in interleaved rounds and the lowest median step wins: here the
    search's pick (one fast step, slow others) loses to a steadier rule whose
    best step did not beat it."""
    import autotune
    times = {8: [1.00, 1.01, 1.02, 1.00, 1.01, 1.02],
             4: [0.80, 0.99, 0.99],   # every step < 1.00: the search's pick
             2: [0.90, 0.91, 0.92]}   # overlaps 0.80 in the search; median 0.91
    dist = np.zeros(3, dtype=np.int32)
    calls = []

    def fake_steps(ctx, xs, ys, rule, reps):
        calls.append(rule["max_cpl"])
        return list(times[rule["max_cpl"]]), dist

    monkeypatch.setattr(autotune, "steps", fake_steps)
    monkeypatch.setattr(autotune, "apply", lambda ctx, rule: None)
    monkeypatch.setattr(autotune, "FINE_SPAN", 0)
    confirmed = []
    best, results, _ = autotune.tune(None, None, None, stages=[{"max_cpl": [4, 2]}],
                                     log=lambda m: None, confirm_top=3, confirmed=confirmed)
    assert best["max_cpl"] == 2
    # finalists: the search's pick first, then by best step (2, then the default)
    assert [r["max_cpl"] for r, _ in confirmed] == [4, 2, 8]
    assert calls[3:] == [4, 2, 8] * autotune.CONFIRM_ROUNDS


def test_autotune_confirmation_averages_workloads(monkeypatch):
    """With a second workload, the confirmation scores each finalist by the
    mean of its median steps on both: a rule that is fast on the tuning
    workload but slow on the second one loses."""
    import autotune
    times = {8: [1.00, 1.01, 1.02, 1.00, 1.01, 1.02], 4: [0.80, 0.99, 0.99], 2: [0.90, 0.91, 0.92]}
    second = {8: [1.0] * 3, 4: [0.9] * 3, 2: [1.2] * 3}
    dist = np.zeros(3, dtype=np.int32)

    def fake_steps(ctx, xs, ys, rule, reps):
        table = second if xs == "x2" else times
        return list(table[rule["max_cpl"]]), dist

    monkeypatch.setattr(autotune, "steps", fake_steps)
    monkeypatch.setattr(autotune, "apply", lambda ctx, rule: None)
    monkeypatch.setattr(autotune, "FINE_SPAN", 0)
    confirmed = []
    best, _, _ = autotune.tune(None, None, None, stages=[{"max_cpl": [4, 2]}], log=lambda m: None,
                               confirm_top=3, confirmed=confirmed, confirm_sets=[("x2", "y2")])
    assert best["max_cpl"] == 4
    scores = {r["max_cpl"]: sc for r, sc in confirmed}
    assert scores[4] == pytest.approx((0.99 + 0.9) / 2)
    assert scores[2] == pytest.approx((0.91 + 1.2) / 2)
