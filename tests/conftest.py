import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libap26gpu.so)")
    config.addinivalue_line("markers", "slow: minutes of CPU; opt in with AP26_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("AP26_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle run; set AP26_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def load_golden_index():
    with open(os.path.join(GOLDEN, "ref_min10", "index.json")) as f:
        return json.load(f)


def read_sol(path):
    """(sorted records [(k, K, first)], hash line) of a SOL-AP26.txt."""
    lines = open(path).read().splitlines()
    recs = sorted(tuple(int(x) for x in l.split()) for l in lines[:-1])
    return recs, lines[-1]


@pytest.fixture(scope="session")
def golden_index():
    return load_golden_index()


@pytest.fixture(scope="session")
def gpu_ctx():
    from ap26_amd import gpu
    ctx = gpu.Context(0)
    yield ctx
    ctx.close()
