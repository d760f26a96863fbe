import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG_DIR = os.path.join(ROOT, "parallel-strong-connectivity_amd")


def load_psc():
    """Import the product package (its directory name is not an identifier)."""
    if "psc_amd" in sys.modules:
        return sys.modules["psc_amd"]
    spec = importlib.util.spec_from_file_location(
        "psc_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["psc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def psc():
    return load_psc()


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    return o


def random_digraph(psc, n, m, seed, self_loops=False, dups=False):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    if not dups:
        key = np.unique(src.astype(np.int64) * n + dst)
        src, dst = key // n, key % n
    return psc.Graph.from_edges(n, src, dst)
