import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libptmf.so)")


def gpu_available() -> bool:
    try:
        from ptatin3d_amd import _lib

        return _lib.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a gpu-marked test on a box without a GPU is an error."""
    from ptatin3d_amd import _lib

    n = _lib.device_count()
    assert n > 0, "no GPU visible to libptmf.so"
    return n


def rel_inf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# -- sub-boxes of a structured Q2 box (node-major [3N], element-major data) --

def sub_slices(m, lo, n):
    """Node slices (z, y, x) of the sub-box of n elements at element lo."""
    return (slice(2 * lo[2], 2 * (lo[2] + n[2]) + 1), slice(2 * lo[1], 2 * (lo[1] + n[1]) + 1),
            slice(2 * lo[0], 2 * (lo[0] + n[0]) + 1))


def sub_nodes(m, a, lo, n):
    X, Y, Z = (2 * d + 1 for d in m)
    return a.reshape(Z, Y, X, 3)[sub_slices(m, lo, n)]


def sub_elems(m, a, lo, n, per):
    return a.reshape(m[2], m[1], m[0], per)[lo[2]:lo[2] + n[2], lo[1]:lo[1] + n[1], lo[0]:lo[0] + n[0]]


def complete_nodes(m, lo, n):
    """Mask of the sub-box nodes whose full-size sums are complete."""
    keep = []
    for d in (2, 1, 0):
        k = np.zeros(2 * n[d] + 1, bool)
        k[1:-1] = True
        k[0] = lo[d] == 0
        k[-1] = lo[d] + n[d] == m[d]
        keep.append(k)
    return keep[0][:, None, None] & keep[1][None, :, None] & keep[2][None, None, :]
