import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def objectives():
    with open(os.path.join(GOLDEN, "objectives.json")) as f:
        return json.load(f)["entries"]


def netlib_path(name):
    return os.path.join(GOLDEN, "netlib", name + ".mps")


def tiny_lp():
    """min x1 + 2 x2 + 0 x3  s.t.  x1 + x2 + x3 = 1,  x1 - x2 <= 0.5,
    0 <= x <= 10.  Optimum x = (0, 0, 1), objective 0."""
    from igaos_amd.lpgen import LP, INF
    rows = [0, 0, 0, 1, 1]
    cols = [0, 1, 2, 0, 1]
    vals = [1.0, 1.0, 1.0, 1.0, -1.0]
    return LP.from_triplets(2, 3, rows, cols, vals, [1.0, 2.0, 0.0],
                            [0.0] * 3, [10.0] * 3, [1.0, -INF], [1.0, 0.5],
                            name="tiny")
