import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "mhd-gpu_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmhdgpu.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "golden_ref.npz")))


@pytest.fixture(scope="session")
def fixture16(oracle):
    rec = np.fromfile(os.path.join(GOLDEN, "source_init_alfvenlinear_16cube.dat"), dtype="<f4")
    return oracle.read_ref_order(rec, 16)


@pytest.fixture(scope="session")
def fixture32(oracle):
    rec = np.fromfile(os.path.join(GOLDEN, "source_init_alfvenlinear_32cube.dat"), dtype="<f4")
    return oracle.read_ref_order(rec, 32)


def same(a, b):
    """Numerically identical (bit for bit except the sign of exact zeros)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def bitwise(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()
