"""Writes the golden fixtures under tests/golden/ (run in a container that
has the reference sources and the compiled reference, oracle/_ref/).

1. ref_pairs.npz: seeded tree pairs (inputs, the generator call that made
   them, distance and full tree-distance table) evaluated by the reference's
   own X-TED_CPU compiled from its sources (oracle/_ref/libxted_ref.so).
2. datasets/: trees 0 and 1 of sampled reference dataset files (data lines
   copied verbatim) and expected.json with the distance the compiled reference
   command line (oracle/_ref/X-TED_CPU) prints for them.

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "x-ted_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402
import xted  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# (name, kind, n1, n2, alphabet, seed config)
PAIRS = [
    ("rrt100_a50_s0", 0, 100, 100, 50, 0),
    ("rrt100_a50_s1", 0, 100, 100, 50, 1),
    ("rrt100_a50_s2", 0, 100, 100, 50, 2),
    ("deep120_a20", 1, 120, 90, 20, 3),
    ("shallow150_a30", 2, 150, 110, 30, 4),
    ("rrt60_a3", 0, 60, 75, 3, 5),
    ("mixed_deep_shallow", 3, 130, 140, 40, 6),
]

# (fixture name, dataset file relative to the reference, first line index)
DATASETS = [
    ("swissport_100", "Sampled_Dataset/1_Swissport/swissport_nodes_100.txt",
     "Sampled_Dataset/1_Swissport/swissport_nodes_adj_100.txt", 0),
    ("python_100", "Sampled_Dataset/2_Python/python_nodes_100.txt",
     "Sampled_Dataset/2_Python/python_nodes_adj_100.txt", 0),
    ("python_100_trees_2_3", "Sampled_Dataset/2_Python/python_nodes_100.txt",
     "Sampled_Dataset/2_Python/python_nodes_adj_100.txt", 2),
    ("dblp_100", "Sampled_Dataset/4_DBLP/dblp_nodes_100.txt",
     "Sampled_Dataset/4_DBLP/dblp_nodes_adj_100.txt", 0),
    ("bolzano", "Sampled_Dataset/3_Bolzano/bolzano_nodes.txt",
     "Sampled_Dataset/3_Bolzano/bolzano_nodes_adj.txt", 0),
    ("swissport_1000", "Sampled_Dataset/1_Swissport/swissport_nodes_1000.txt",
     "Sampled_Dataset/1_Swissport/swissport_nodes_adj_1000.txt", 0),
    ("python_1000", "Sampled_Dataset/2_Python/python_nodes_1000.txt",
     "Sampled_Dataset/2_Python/python_nodes_adj_1000.txt", 0),
    ("dblp_1000", "Sampled_Dataset/4_DBLP/dblp_nodes_1000.txt",
     "Sampled_Dataset/4_DBLP/dblp_nodes_adj_1000.txt", 0),
    ("xl_synthetic_1000", "Sampled_Dataset_Extra_Large_Trees/synthetic/1000_nodes/synthetic_nodes_1000.txt",
     "Sampled_Dataset_Extra_Large_Trees/synthetic/1000_nodes/synthetic_nodes_adj_1000.txt", 0),
    ("xl_python_1000", "Sampled_Dataset_Extra_Large_Trees/python/1000_nodes/python_nodes_1000.txt",
     "Sampled_Dataset_Extra_Large_Trees/python/1000_nodes/python_nodes_adj_1000.txt", 0),
    ("xl_swissport_1000", "Sampled_Dataset_Extra_Large_Trees/swissport/1000_nodes/swissport_nodes_1000.txt",
     "Sampled_Dataset_Extra_Large_Trees/swissport/1000_nodes/swissport_nodes_adj_1000.txt", 0),
    ("xl_synthetic_2000", "Sampled_Dataset_Extra_Large_Trees/synthetic/2000_nodes/synthetic_nodes_2000.txt",
     "Sampled_Dataset_Extra_Large_Trees/synthetic/2000_nodes/synthetic_nodes_adj_2000.txt", 0),
]


def gen_pair(kind, n1, n2, alphabet, cfg):
    k1, k2 = (kind, kind) if kind < 3 else (1, 2)
    s1, s2 = xted.derive_seed(100 + cfg, 0), xted.derive_seed(100 + cfg, 1)
    return (k1, n1, alphabet, s1), (k2, n2, alphabet, s2)


def write_pairs():
    out = {}
    for name, kind, n1, n2, alphabet, cfg in PAIRS:
        g1, g2 = gen_pair(kind, n1, n2, alphabet, cfg)
        x = xted.generate(*g1)
        y = xted.generate(*g2)
        d, td, _ = oracle.ref_ted(x, y, threads=2)
        assert td[0, 0] == d
        out[f"{name}/x_parent"], out[f"{name}/x_label"] = x
        out[f"{name}/y_parent"], out[f"{name}/y_label"] = y
        out[f"{name}/gen"] = np.array([g1, g2], dtype=np.uint64)
        out[f"{name}/td"] = td
        out[f"{name}/dist"] = np.array([d], np.int32)
        print(f"{name}: {len(x[0])}x{len(y[0])} dist {d}")
    np.savez_compressed(os.path.join(HERE, "ref_pairs.npz"), **out)


def write_datasets(ref_root):
    expected = {}
    for name, labels, adj, first in DATASETS:
        with open(os.path.join(ref_root, labels)) as f:
            llines = f.read().split("\n")
        with open(os.path.join(ref_root, adj)) as f:
            alines = f.read().split("\n")
        lp = os.path.join(HERE, "datasets", f"{name}_nodes.txt")
        ap = os.path.join(HERE, "datasets", f"{name}_adj.txt")
        with open(lp, "w") as f:
            f.write(llines[first] + "\n" + llines[first + 1] + "\n")
        with open(ap, "w") as f:
            f.write(alines[first] + "\n" + alines[first + 1] + "\n")
        # The reference CLI repeats the pair 20 times; all runs must agree.
        out = subprocess.run([oracle.REF_CLI, "1", lp, ap], capture_output=True, text=True,
                             check=True).stdout
        ds = set(int(v) for v in re.findall(r"The distance = (-?\d+)", out))
        assert len(ds) == 1, ds
        expected[name] = ds.pop()
        print(f"{name}: {expected[name]}")
    with open(os.path.join(HERE, "datasets", "expected.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    write_pairs()
    write_datasets(ref_root)
