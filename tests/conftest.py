import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

PKG_DIR = os.path.join(REPO, "boundarylayerdynamics.jl_amd")


def load_package():
    """Import the host package (its directory name is not a Python identifier)."""
    if "bld_amd" in sys.modules:
        return sys.modules["bld_amd"]
    spec = importlib.util.spec_from_file_location(
        "bld_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bld_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")


def convergence(N, eps, exponential=False, order=None, threshold_linearity=0.99, threshold_slope=0.95):
    """test/test_utils.jl:11-43: least-squares slope of log(ε) over log(N) (or N)."""
    X = np.asarray(N, dtype=float) if exponential else np.log(np.asarray(N, dtype=float))
    Y = np.log(np.asarray(eps, dtype=float))
    n = len(X)
    cov = lambda x, y: np.sum((x - x.mean()) * (y - y.mean())) / (n - 1)
    slope = cov(X, Y) / cov(X, X)
    rsq = cov(X, Y) ** 2 / (cov(X, X) * cov(Y, Y))
    assert rsq > threshold_linearity, (rsq, eps)
    if order is not None:
        assert -slope > order * threshold_slope, (-slope, eps)
    return -slope, rsq


def isapprox(a, b, rtol=np.sqrt(np.finfo(float).eps)):
    """Julia's `≈` on arrays: norm(a-b) <= rtol * max(norm(a), norm(b))."""
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) <= rtol * max(np.linalg.norm(a.ravel()), np.linalg.norm(b.ravel()))
