"""The exact coarse solve of any size (coarse_factor 2: multifrontal Cholesky
of the assembled level on a nested-dissection ordering) through the C-ABI.

A one-level hierarchy's V-cycle is the coarse solve itself, x = A^-1 b.  It is
checked against the oracle's operator (true FP64 residual, 1e-10), against
the dense explicit inverse where that exists (1e-10), on the 16^3 coarse level
of the 128^3 workload (107,811 dofs, contrast 1e6), and inside V-cycles
(equal to the dense-coarse V-cycle; graph replays equal to eager cycles).
The reference's coarse solve is PETSc LU on the assembled level
(-fieldsplit_u_mg_coarse_pc_type lu); PETSc is not available here, so the
solve is pinned by the oracle residual, not by LU's iterates."""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray
from ptatin3d_amd.multigrid import GMG, default_options

from conftest import rel_inf

pytestmark = pytest.mark.gpu


def _one_level(m, coords, eta, bc, factor, galerkin=0):
    mg = GMG(*m, 1)
    mg.set_coordinates(coords)
    mg.set_viscosity(0, eta)
    mg.set_dirichlet(0, bc)
    mg.setup(default_options(coarse_factor=factor, smoother_precision=64))
    return mg


def _solve(mg, b):
    db, dx = DeviceArray.from_host(b), DeviceArray(b.size)
    mg.vcycle(db, dx)
    x = dx.download()
    db.free()
    dx.free()
    return x


@pytest.mark.parametrize("m", [(6, 5, 4), (3, 7, 2), (1, 1, 1), (9, 4, 5)])
def test_sparse_coarse_solve_is_exact(gpu, m):
    rng = np.random.default_rng(sum(m))
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    E = m[0] * m[1] * m[2]
    eta = F.eta_loguniform(E, 1e-3, 1e3, rng)
    bc = F.dirichlet_all_faces(*m)
    b = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    print(f"coarse: box {m}", flush=True)
    sp = _one_level(m, coords, eta, bc, 2)
    assert sp.coarse_info()["kind"] == "sparse"
    x = _solve(sp, b)
    r = b - O.a11_matmult(*m, coords, eta, bc, x)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b)
    # Dirichlet rows are identities
    d = bc.astype(bool)
    assert np.array_equal(x[d], b[d])
    # the dense explicit inverse gives the same solution
    de = _one_level(m, coords, eta, bc, 1)
    assert de.coarse_info()["kind"] == "dense"
    assert rel_inf(x, _solve(de, b)) <= 1e-10
    # repeated solves are bit-identical (fixed summation order)
    assert np.array_equal(_solve(sp, b), x)
    sp.destroy()
    de.destroy()


def test_sparse_coarse_solve_mixed_boundary(gpu):
    """Free-slip sides and base with x-face extension (the rift boundary
    mask): partly Dirichlet nodes inside the fronts."""
    m = (6, 4, 4)
    rng = np.random.default_rng(2)
    coords = F.perturb_interior(F.box_coords(*m, hi=(2.0, 2.0, 1.0)), *m, 0.1, rng)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-2, 1e2, rng)
    bc = F.rift_bc(*m)[0]
    b = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    sp = _one_level(m, coords, eta, bc, 2)
    x = _solve(sp, b)
    r = b - O.a11_matmult(*m, coords, eta, bc, x)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b)
    sp.destroy()


def test_sparse_coarse_solve_16cube_contrast_1e6(gpu):
    """The 16^3 coarse level of the 128^3 contrast-1e6 workload (107,811 dofs,
    beyond the dense inverse's 32,768): exact to a 1e-9 FP64 residual."""
    m = (16, 16, 16)
    coords = F.box_coords(*m)
    eta = F.eta_layered_ellipsoids(F.qp_coordinates(coords, *m), rng=np.random.default_rng(7))
    bc = F.dirichlet_all_faces(*m)
    b = F.smooth_body_force(coords, *m)
    mg = GMG(*m, 1)
    mg.set_coordinates(coords)
    mg.set_viscosity(0, eta)
    mg.set_dirichlet(0, bc)
    mg.setup(default_options(smoother_precision=64))
    info = mg.coarse_info()
    assert info["kind"] == "sparse", info  # chosen automatically above 32,768 dofs
    x = _solve(mg, b)
    r = b - O.a11_matmult(*m, coords, eta, bc, x)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b), np.linalg.norm(r) / np.linalg.norm(b)
    mg.destroy()


@pytest.mark.parametrize("galerkin", [0, 1])
def test_vcycle_with_sparse_coarse_equals_dense(gpu, galerkin):
    """8^3 -> 4^3 V-cycles (FP64, rediscretised or Galerkin coarse level):
    the sparse and the dense coarse factor give the same cycle; the graph
    replays of the sparse one equal its eager cycle."""
    m = (8, 8, 8)
    rng = np.random.default_rng(5)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    eta = F.eta_loguniform(512, 1e-2, 1e2, rng)
    out = {}
    for fac, graph in ((1, 0), (2, 0), (2, 1)):
        mg = GMG(*m, 2)
        mg.set_coordinates(coords)
        mg.set_viscosity(1, eta)
        mg.coarsen_viscosity()
        mg.set_dirichlet_all_faces()
        mg.setup(default_options(coarse_factor=fac, smoother_precision=64, use_graph=graph,
                                 galerkin_levels=galerkin))
        xs = []
        for k in range(4):
            b = np.random.default_rng(100 + k).uniform(-1, 1, 3 * F.nnodes(*m))
            xs.append(_solve(mg, b))
        out[(fac, graph)] = xs
        mg.destroy()
    for k in range(4):
        assert rel_inf(out[(2, 0)][k], out[(1, 0)][k]) <= 1e-10, k
        assert np.array_equal(out[(2, 1)][k], out[(2, 0)][k]), k
