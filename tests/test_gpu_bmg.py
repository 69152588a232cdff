"""GPU tests of the batched multigrid (many boxes in lockstep) through the
C-ABI: the batched V-cycle equals each box's own V-cycle (also on graph
replays), and the batched flexible CG drives every box to its tolerance,
checked with the oracle's operator on the host.

Tolerances: V-cycle FP64 1e-10, FP32 1e-5 relative (the batched coarse
product sums in another order than the single-box one); solve: true FP64
residual <= 2e-8 ||b||.  Log-uniform viscosity in [1e-3, 1e3] per quadrature
point and box sides up to 24 elements, as the batched workload.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray
from ptatin3d_amd.multigrid import GMG, BatchGMG, default_options

from conftest import rel_inf

pytestmark = pytest.mark.gpu

# sides from the batched workload's {4, 8, ..., 24}, one box 24 elements long
DIMS = [(4, 4, 4), (8, 4, 8), (12, 8, 4), (8, 8, 8), (16, 8, 12), (4, 12, 8), (24, 8, 4)]


def _boxes(prec, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for m in DIMS:
        coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
        eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-3, 1e3, rng)  # the workload's contrast 1e6
        mg = GMG(*m, 3)
        mg.set_coordinates(coords)
        mg.set_viscosity(2, eta)
        mg.coarsen_viscosity("qp-arithmetic")
        mg.set_dirichlet_all_faces()
        mg.setup(default_options(smoother_precision=prec, pre_its=3, post_its=3, use_graph=0))
        out.append((m, coords, eta, mg))
    return out, rng


@pytest.mark.parametrize("prec", [64, 32])
def test_batched_vcycle_equals_per_box(gpu, prec):
    boxes, rng = _boxes(prec)
    mgs = [b[3] for b in boxes]
    bs = [DeviceArray.from_host(rng.uniform(-1, 1, mg.n)) for mg in mgs]
    ref = []
    for mg, b in zip(mgs, bs):
        x = DeviceArray(mg.n)
        mg.vcycle(b, x)
        ref.append(x.download())
    bm = BatchGMG(mgs)
    xs = [DeviceArray(mg.n) for mg in mgs]
    tol = 1e-10 if prec == 64 else 1e-5
    for rep in range(3):  # eager, capture, replay
        for x in xs:
            x.zero()
        bm.vcycle(bs, xs)
        for i, (x, r) in enumerate(zip(xs, ref)):
            assert rel_inf(x.download(), r) <= tol, (rep, i)
    bm.destroy()
    for mg in mgs:
        mg.destroy()


def test_batched_fcg_solves_every_box(gpu):
    boxes, rng = _boxes(32, seed=9)
    mgs = [b[3] for b in boxes]
    bh = [F.random_velocity(*m, F.dirichlet_all_faces(*m), rng) for (m, _, _, _) in boxes]
    bs = [DeviceArray.from_host(b) for b in bh]
    xs = [DeviceArray(mg.n) for mg in mgs]
    for x in xs:
        x.zero()
    bm = BatchGMG(mgs)
    its, rn = bm.fcg(bs, xs, rtol=1e-8, maxit=300)
    for i, ((m, coords, eta, mg), b, x) in enumerate(zip(boxes, bh, xs)):
        assert its[i] > 0
        u = x.download()
        r = b - O.a11_matmult(*m, coords, eta, F.dirichlet_all_faces(*m), u)
        assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(b), (i, its[i])
        assert rn[i] <= 2e-8 * np.linalg.norm(b)
        # same iteration count as the box's own flexible CG, within a few
        y = DeviceArray(mg.n)
        y.zero()
        n1, _, _ = mg.fcg(bs[i], y, rtol=1e-8, maxit=300)
        assert abs(n1 - its[i]) <= 3, (i, n1, its[i])
    bm.destroy()
    for mg in mgs:
        mg.destroy()

