"""Shared fixtures. `gpu` tests need an MI355X; everything else runs on the CPU."""
import glob
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def tbem():
    return importlib.import_module("3bem_amd")


@pytest.fixture(scope="session")
def wl():
    return importlib.import_module("3bem_amd.workloads")


@pytest.fixture(scope="session")
def orc():
    from oracle.oracle import Oracle
    return Oracle("orc")


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import Oracle, available
    if not available("ref"):
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    return Oracle("ref")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def elastic_rows():
    """Rows of the reference's sympy golden data test_data_elastic3
    (columns per unit_tests/test_elastic_kernels.cpp:54-103)."""
    rows = []
    with open(os.path.join(GOLDEN, "test_data_elastic3")) as f:
        for line in f:
            es = line.strip().split(",")
            if len(es) < 20:
                continue
            rows.append(dict(value=float(es[0]), name=es[1], k=int(es[2]), j=int(es[3]),
                             mu=float(es[4]), nu=float(es[5]),
                             src=np.array([float(v) for v in es[8:11]]),
                             obs=np.array([float(v) for v in es[11:14]]),
                             nsrc=np.array([float(v) for v in es[14:17]]),
                             nobs=np.array([float(v) for v in es[17:20]])))
    return rows


ELASTIC_NAMES = {"disp": 3, "trac": 4, "adj_trac": 5, "hyp": 6}
