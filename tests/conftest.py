import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("x-ted_amd", "oracle"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def load_dataset_pair(name):
    """Trees 0 and 1 of a golden dataset fixture as (parent, label) with the
    labels interned by string equality (the reference compares strings)."""
    import xted_text
    return xted_text.load_pair(os.path.join(GOLDEN, "datasets", f"{name}_nodes.txt"),
                               os.path.join(GOLDEN, "datasets", f"{name}_adj.txt"))


def golden_pairs():
    data = np.load(os.path.join(GOLDEN, "ref_pairs.npz"))
    names = sorted({k.split("/")[0] for k in data.files})
    out = []
    for nm in names:
        out.append((nm, (data[f"{nm}/x_parent"], data[f"{nm}/x_label"]),
                    (data[f"{nm}/y_parent"], data[f"{nm}/y_label"]), data[f"{nm}/td"],
                    int(data[f"{nm}/dist"][0]), data[f"{nm}/gen"]))
    return out


@pytest.fixture(scope="session")
def gpu_ctx():
    import xted
    ctx = xted.Context(0)
    yield ctx
    ctx.close()
