import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_fixture(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    d["params"] = json.loads(str(d["params"]))
    return d


def rel_err(a, b):
    """max |a-b| / max |b| (per-field scale)"""
    s = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (s if s > 0 else 1.0)


def field_errors(a, b):
    return [rel_err(a[k], b[k]) for k in range(a.shape[0])]


GROUPS = {"velocity": [0, 1], "stress": [2, 3, 4, 5], "chi": [6], "refmap": [7, 8], "ddev": [9]}


def group_errors(a, b):
    """max |a-b| / max |b| over each physical group of planes (u,v share one
    scale, the four stress components another)"""
    return {g: rel_err(a[ix], b[ix]) for g, ix in GROUPS.items()}


@pytest.fixture(scope="session")
def gpu_available():
    import stz_amd
    try:
        return stz_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _skip_gpu_without_device(request):
    if request.node.get_closest_marker("gpu") is not None:
        import stz_amd
        if stz_amd.device_count() == 0:
            pytest.skip("no GPU device visible")
