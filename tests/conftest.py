import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "jianhui-lbm_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import the package directory (its name has a hyphen) as ``jianhui_lbm_amd``."""
    if "jianhui_lbm_amd" in sys.modules:
        return sys.modules["jianhui_lbm_amd"]
    spec = importlib.util.spec_from_file_location("jianhui_lbm_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["jianhui_lbm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


def load_fixture(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    params = json.loads(str(d["params"]))
    return d, params


def rel_err(a, b):
    """Norm-wise relative error ||a-b|| / ||b||."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
