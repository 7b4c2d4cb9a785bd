"""Reference text formats in Python: label lines and preorder adjacency lines
(README.md:31-44 of the reference; tokenisation of
X-TED_GPU/dataset_processing.cu:7-82), with labels interned to int32 ids by
string equality."""
import numpy as np


def split_labels(line):
    """A token is emitted only at a space (a trailing token without a space is
    dropped, as the reference's node_process does)."""
    out, part = [], []
    for ch in line:
        if ch != " ":
            part.append(ch)
        else:
            out.append("".join(part))
            part = []
    return out


def parse_adjacency(line):
    body = line[1:-1]
    adj, i = [], 0
    while True:
        i = body.find("[", i)
        if i < 0:
            break
        j = body.find("]", i)
        if j < 0:
            raise ValueError("unterminated '[' group")
        inner = body[i + 1:j].replace(" ", "")
        adj.append([int(t) for t in inner.split(",")] if inner else [])
        i = j + 1
    return adj


def adj_to_parent(adj):
    n = len(adj)
    parent = np.full(n, -2, np.int32)
    if n:
        parent[0] = -1
    for i, ch in enumerate(adj):
        for c in ch:
            if not 0 < c < n or parent[c] != -2:
                raise ValueError("not a tree")
            parent[c] = i
    if (parent == -2).any():
        raise ValueError("not a tree")
    return parent


def load_trees(labels_path, adj_path, ids=None):
    """Every tree of a label file and its adjacency file, labels interned
    through `ids` (a dict shared across files, so equal strings get equal
    ids). Blank trailing lines are skipped."""
    if ids is None:
        ids = {}
    with open(labels_path) as f:
        llines = f.read().split("\n")
    with open(adj_path) as f:
        alines = f.read().split("\n")
    trees = []
    for lline, aline in zip(llines, alines):
        if not aline.strip():
            continue
        adj = parse_adjacency(aline)
        parent = adj_to_parent(adj)
        labels = split_labels(lline)
        lab = np.array([ids.setdefault(s, len(ids)) for s in labels[:len(adj)]], np.int32)
        trees.append((parent, lab))
    return trees


def load_pair(labels_path, adj_path, first=0):
    with open(labels_path) as f:
        llines = f.read().split("\n")
    with open(adj_path) as f:
        alines = f.read().split("\n")
    ids = {}
    trees = []
    for k in (first, first + 1):
        labels = split_labels(llines[k])
        adj = parse_adjacency(alines[k])
        parent = adj_to_parent(adj)
        lab = np.array([ids.setdefault(s, len(ids)) for s in labels[:len(adj)]], np.int32)
        trees.append((parent, lab))
    return trees[0], trees[1]
