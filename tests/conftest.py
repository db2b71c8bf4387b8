import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "unifrac-binaries_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs through libssu.so on the device)")
    # the checker and the library are built ahead of time; build them here only if absent
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle"], check=True)
    if not os.path.exists(os.path.join(PKG, "libssu.so")):
        subprocess.run(["make", "-s", "-C", PKG, "-j8"], check=True)


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def ssu():
    import ssu as m
    return m


@pytest.fixture(scope="session")
def synth():
    import synth as m
    return m


@pytest.fixture(scope="session")
def goldens():
    return load_json("test_su_goldens.json")


@pytest.fixture(scope="session")
def ref_tables():
    return load_json("tables.json")


@pytest.fixture(scope="session")
def ref_trees():
    return load_json("trees.json")


@pytest.fixture(scope="session")
def ref_expected():
    return load_json("expected.json")


def stripes_of(matrix, n_stripes=None):
    """Stripe layout of a full matrix: row s, element k = M[k][(k+s+1) % n]."""
    m = np.asarray(matrix)
    n = m.shape[0]
    if n_stripes is None:
        n_stripes = (n + 1) // 2
    k = np.arange(n)
    return np.stack([m[k, (k + s + 1) % n] for s in range(n_stripes)])


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    both_nan = np.isnan(a) & np.isnan(b)
    scale = np.maximum(np.abs(b), floor)
    r = np.where(b == 0, d, d / scale)
    r = np.where(both_nan, 0.0, r)
    return float(np.nanmax(np.where(np.isnan(a) ^ np.isnan(b), np.inf, r))) if r.size else 0.0
