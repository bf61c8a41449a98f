import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "webgraph-algo-rs_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")
    hb = _load("hyperball", os.path.join(PKG, "hyperball.py"))
    orc = _load("hb_oracle_py", os.path.join(REPO, "oracle", "oracle.py"))
    # Build only when nothing is built yet.
    if not os.path.exists(hb.LIB_PATH) or not os.path.exists(orc.LIB_PATH):
        _load("graft_entry", os.path.join(REPO, "__graft_entry__.py")).build()


@pytest.fixture(scope="session")
def hbm():
    return sys.modules["hyperball"]


@pytest.fixture(scope="session")
def orc():
    return sys.modules["hb_oracle_py"]


@pytest.fixture(scope="session")
def cnr(hbm):
    base = os.path.join(GOLDEN, "cnr-2000", "cnr-2000")
    return hbm.Graph.load_bvgraph(base), hbm.Graph.load_bvgraph(base + "-t")
