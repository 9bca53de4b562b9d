import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"),
          os.path.join(ROOT, "fortran-cluster-md_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    """(variant, case dict, actions, list of record dicts) from a fixture."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    case = {}
    recs = {}
    for k in z.files:
        if k.startswith("in_"):
            v = z[k]
            case[k[3:]] = v.item() if v.ndim == 0 else v
        elif k[0] == "r" and "_" in k:
            idx, key = k[1:].split("_", 1)
            v = z[k]
            recs.setdefault(int(idx), {})[key] = v.item() if v.ndim == 0 else v
    return str(z["variant"]), case, [tuple(a) for a in z["actions"]], [recs[i] for i in sorted(recs)]


@pytest.fixture
def golden():
    return load_golden
