"""Batched MatMult of boxes of different sizes equals the per-box MatMult
(and the oracle), in FP64 and FP32, through the C-ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray
from ptatin3d_amd.operators import BatchA11, MatA11MF

from conftest import rel_inf

pytestmark = pytest.mark.gpu

SIZES = [(4, 4, 4), (8, 4, 12), (5, 7, 3), (12, 12, 12), (1, 2, 1), (16, 8, 8), (3, 9, 4)]


def _boxes(seed=23):
    rng = np.random.default_rng(seed)
    out = []
    for m in SIZES:
        coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
        eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-3, 1e3, rng)
        bc = F.dirichlet_all_faces(*m)
        x = rng.uniform(-1, 1, 3 * F.nnodes(*m))
        A = MatA11MF(*m).setup(coords, eta, bc)
        out.append((m, coords, eta, bc, x, A))
    return out


def test_batch_matches_single_and_oracle(gpu):
    boxes = _boxes()
    # give the boxes different tilings too
    boxes[3][5].set_tiling(7, 3, 2)
    boxes[5][5].set_tiling(4, 5, 1)
    B = BatchA11([b[5] for b in boxes])
    for prec, dt, tol in ((64, np.float64, 1e-14), (32, np.float32, 1e-6)):
        xs = [DeviceArray.from_host(b[4].astype(dt)) for b in boxes]
        ys = [DeviceArray(b[4].size, dt) for b in boxes]
        for rep in range(2):
            B.mult(xs, ys, precision=prec)
            for (m, coords, eta, bc, x, A), y in zip(boxes, ys):
                yb = y.download()
                ys_single = A.mult_host(x, precision=prec)
                assert rel_inf(yb, ys_single) <= tol, (m, prec)
                if prec == 64:
                    assert rel_inf(yb, O.a11_matmult(*m, coords, eta, bc, x)) <= 1e-12, m
    B.destroy()
    for b in boxes:
        b[5].destroy()
