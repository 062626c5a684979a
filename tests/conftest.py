import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def read_xyz(path):
    """Reader of the reference's XYZ input format (SRC/SphereSystem.cpp:129-156):
    two header lines, then rows `type gid radius x y z`."""
    rows = []
    with open(path) as f:
        f.readline()
        f.readline()
        for line in f:
            t = line.split()
            if len(t) >= 6:
                rows.append((int(t[1]), float(t[2]), float(t[3]), float(t[4]), float(t[5])))
    a = np.array(rows)
    return a[:, 2:5].copy(), a[:, 1].copy()


@pytest.fixture(scope="session")
def pos_initial():
    return read_xyz(os.path.join(GOLDEN, "posInitial.txt"))


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def relmax(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
