import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def mmf():
    import minimmerflow_amd
    minimmerflow_amd.lib()
    return minimmerflow_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


def _border_index(n, dim, owner, normal):
    """Entry of each border face (owner cell, owner-outward normal) in the
    per-face border layout: sides x-, x+, y-, y+ (, z-, z+) one after
    another, a side's faces ordered by its two other axes, the lower one
    fastest."""
    import numpy as np
    n = list(n) + [1] * (3 - len(n))
    owner = np.asarray(owner)
    normal = np.asarray(normal).reshape(-1, 3)
    q = np.stack([owner % n[0], (owner // n[0]) % n[1], owner // (n[0] * n[1])], axis=1)
    a = np.argmax(np.abs(normal), axis=1)
    side = 2 * a + (normal[np.arange(len(a)), a] > 0)
    count = [n[1] * n[2], n[1] * n[2], n[0] * n[2], n[0] * n[2], n[0] * n[1], n[0] * n[1]]
    off = np.concatenate([[0], np.cumsum(count[:2 * dim])])
    b0 = np.where(a == 0, 1, 0)
    b1 = np.where(a == 2, 1, 2)
    rows = np.arange(len(a))
    nb0 = np.array(n)[b0]
    return off[side] + q[rows, b0] + nb0 * q[rows, b1]


@pytest.fixture(scope="session")
def border_index():
    return _border_index
