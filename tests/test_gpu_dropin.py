"""The reference-signature drop-in run as a drop-in: a C++ caller written
against include/TED_C++.h (tests/dropin_check.cpp, built by
x-ted_amd/Makefile) calls standard_ted, standard_ted_1,
outermost_right_leaves, key_roots, parallel_standard_ted_test and min3 on the
reference's own dataset pairs, and the X-TED_GPU command line runs as
X-TED_GPU/main.cu:4-14 does. Results are compared with the dataset known
answers (tests/golden/datasets/expected.json, written by the compiled
reference) and with the oracle's full tables."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT, load_dataset_pair

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "x-ted_amd", "bin")
DATA = os.path.join(GOLDEN, "datasets")
with open(os.path.join(DATA, "expected.json")) as _f:
    EXPECTED = json.load(_f)
# Pairs whose full tables are also compared with the oracle's (the larger
# ones are checked by distance, orl and keyroots).
FULL_TABLES = {"bolzano", "dblp_100", "python_100", "swissport_100", "python_100_trees_2_3",
               "swissport_1000", "dblp_1000"}


def files(name):
    return os.path.join(DATA, f"{name}_nodes.txt"), os.path.join(DATA, f"{name}_adj.txt")


def quiet_env():
    env = dict(os.environ)
    env["XTED_QUIET"] = "1"
    return env


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_cpp_dropin_entry_points(name, tmp_path):
    exe = os.path.join(BIN, "dropin_check")
    assert os.path.exists(exe), "build x-ted_amd first (make -C x-ted_amd)"
    prefix = str(tmp_path / name)
    print(f"dropin_check {name}", flush=True)
    r = subprocess.run([exe, *files(name), prefix], capture_output=True, text=True, timeout=120,
                       env=quiet_env())
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    m, n = out["m"], out["n"]
    x, y = load_dataset_pair(name)
    assert (m, n) == (len(x[0]), len(y[0]))
    assert out["min3"] == [1, -4, 6]
    # standard_ted: the known answer, and the caller's label vectors cleared
    # (X-TED_GPU/TED.cu:93-94), for standard_ted and standard_ted_1 alike.
    assert out["standard_ted"] == EXPECTED[name]
    assert out["standard_ted_cleared"] and out["standard_ted_1_cleared"]
    assert out["standard_ted_1_shape"] == [m, n]
    assert out["parallel_standard_ted_test_shape"] == [m, n]
    # empty-tree rule (X-TED_GPU/TED.cu:119-124): the other tree's size
    assert out["empty_x"] == m and out["empty_y"] == m
    # orl and keyroots against the oracle (closed forms pinned there)
    assert out["x_orl"] == oracle.orl(x[0]).tolist()
    assert out["y_orl"] == oracle.orl(y[0]).tolist()
    assert out["x_kr"] == oracle.keyroots(oracle.orl(x[0])).tolist()
    assert out["y_kr"] == oracle.keyroots(oracle.orl(y[0])).tolist()
    td1 = np.fromfile(prefix + "_standard_ted_1.bin", np.int32).reshape(m, n)
    td2 = np.fromfile(prefix + "_parallel_standard_ted_test.bin", np.int32).reshape(m, n)
    assert td1[0, 0] == EXPECTED[name]
    # The reference's GPU scheduler returns a zero matrix
    # (X-TED_GPU/parallel_X_TED.cu:393-397); this one returns the table.
    assert (td2 == td1).all()
    if name in FULL_TABLES:
        _, otd, _ = oracle.ted(x, y, threads=8)
        assert (td1 == otd).all(), f"{int((td1 != otd).sum())} cells differ"


@pytest.mark.parametrize("name", ["swissport_100", "python_1000", "xl_synthetic_2000"])
def test_command_line_prints_twenty_distances(name):
    """X-TED_GPU <labels> <adj>: 20 runs of test() (X-TED_GPU/main.cu:4-14),
    each printing the distance of trees 0 and 1 and a blank line."""
    print(f"X-TED_GPU {name}", flush=True)
    r = subprocess.run([os.path.join(BIN, "X-TED_GPU"), *files(name)], capture_output=True,
                       text=True, timeout=120, env=quiet_env())
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    dist = [ln for ln in lines if ln.startswith("The distance = ")]
    assert dist == [f"The distance = {EXPECTED[name]}"] * 20
    assert lines.count("") >= 20
    # with the timing lines on, each run also prints the reference's two
    # timing lines (X-TED_GPU/parallel_X_TED.cu:224,392)
    r2 = subprocess.run([os.path.join(BIN, "X-TED_GPU"), *files(name)], capture_output=True,
                        text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout.count("Measured time for total parallel execution = ") == 20
    assert r2.stdout.count("Preprocessing Method, the time is ") == 20
