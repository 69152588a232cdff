"""GPU tests of the saddle-point pieces against the CPU oracle, through the C-ABI.

The Schur blocks and the saddle-point operator are compared with the oracle
(1e-12 relative); the FGMRES solve with the block-triangular preconditioner
is checked through its true FP64 residual, recomputed with the oracle's
A11, A12 and A21.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray
from ptatin3d_amd.multigrid import GMG, default_options, stokes_solve_timed
from ptatin3d_amd.operators import MatStokesMF

from conftest import rel_inf

pytestmark = pytest.mark.gpu


def _case(m, seed=11, lo=1e-2, hi=1e2):
    rng = np.random.default_rng(seed)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], lo, hi, rng)
    return coords, eta, F.dirichlet_all_faces(*m), rng


def test_schur_blocks_match_oracle(gpu):
    m = (5, 4, 3)
    coords, eta, bc, _ = _case(m)
    A = MatStokesMF(*m).setup(coords, eta, bc)
    assert rel_inf(A.schur_blocks_host().reshape(-1), O.schur_mass_blocks(*m, coords, eta).reshape(-1)) <= 1e-12
    A.destroy()


def _oracle_stokes(m, coords, eta, bc, u, p):
    yu = O.a11_matmult(*m, coords, eta, bc, u) + O.a12_matmult(*m, coords, bc, p)
    return yu, O.a21_matmult(*m, coords, bc, u)


def _hier(m, nl, coords, eta, prec=64):
    mg = GMG(*m, nl)
    mg.set_coordinates(coords)
    mg.set_viscosity(nl - 1, eta)
    mg.coarsen_viscosity()
    mg.set_dirichlet_all_faces()
    mg.setup(default_options(smoother_precision=prec))
    mg.stokes_setup()
    return mg


def test_stokes_operator_matches_oracle(gpu):
    m = (4, 4, 4)
    coords, eta, bc, rng = _case(m)
    mg = _hier(m, 2, coords, eta)
    u = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    p = rng.uniform(-1, 1, 4 * m[0] * m[1] * m[2])
    x = DeviceArray.from_host(np.concatenate([u, p]))
    y = DeviceArray(x.n)
    mg.stokes_apply(x, y)
    yh = y.download()
    yu, yp = _oracle_stokes(m, coords, eta, bc, u, p)
    assert rel_inf(yh[: u.size], yu) <= 1e-12
    assert rel_inf(yh[u.size:], yp) <= 1e-12
    mg.destroy()


@pytest.mark.parametrize("prec", [64, 32])
def test_stokes_fgmres_converges(gpu, prec):
    m = (8, 8, 8)
    coords, eta, bc, rng = _case(m)
    mg = _hier(m, 3, coords, eta, prec)
    bu = F.smooth_body_force(coords, *m)
    bp = np.zeros(4 * m[0] * m[1] * m[2])
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, bp, rtol=1e-8, maxit=400, restart=60)
    yu, yp = _oracle_stokes(m, coords, eta, bc, u, p)
    r = np.concatenate([bu - yu, bp - yp])
    assert np.linalg.norm(r) <= 1.5e-8 * np.linalg.norm(bu), (its, np.linalg.norm(r) / np.linalg.norm(bu))
    mg.destroy()


@pytest.mark.parametrize("m,tiling", [((5, 4, 3), None), ((1, 1, 1), None), ((22, 1, 3), None),
                                      ((8, 9, 11), (4, 5, 2)), ((7, 6, 5), (1, 1, 1)), ((9, 7, 8), (7, 3, 3))])
def test_fused_stokes_mult_matches_oracle(gpu, m, tiling):
    """MatMult_MFStokes_A in one pass = oracle A11 + A12 and A21."""
    coords, eta, bc, rng = _case(m)
    A = MatStokesMF(*m).setup(coords, eta, bc)
    if tiling:
        A.set_tiling(*tiling)
    u = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    p = rng.uniform(-1, 1, 4 * m[0] * m[1] * m[2])
    yu, yp = A.stokes_host(u, p)
    ou, op = _oracle_stokes(m, coords, eta, bc, u, p)
    assert rel_inf(yu, ou) <= 1e-12
    assert rel_inf(yp, op) <= 1e-12
    # scaled elements far from the origin: the centroid-relative pressure
    # basis keeps its accuracy
    c2 = 1e3 + 10.0 * coords
    A.set_coordinates(c2)
    yu, yp = A.stokes_host(u, p)
    ou, op = _oracle_stokes(m, c2, eta, bc, u, p)
    assert rel_inf(yu, ou) <= 1e-12
    assert rel_inf(yp, op) <= 1e-12
    A.destroy()


def test_rift_saddle_point_with_boundary_values(gpu):
    """Rift workload (free-slip sides and base, extension on the x-faces, free
    top, power-law viscosity of a seeded velocity, contrast 1e6) at 16x16x8:
    the lifted right-hand side gives a solution that carries the boundary
    values and satisfies the unconstrained equations at the free dofs, all
    recomputed with the oracle's operators."""
    from ptatin3d_amd.multigrid import powerlaw_viscosity, rift_hierarchy, stokes_rhs_lifted

    dims = (16, 16, 8)
    c = F.config5(*dims)
    eta = powerlaw_viscosity(dims, c["coords"], c["u0"], c["B"], c["nexp"], c["eta_min"], c["eta_max"])
    assert rel_inf(eta, O.powerlaw_eta(*dims, c["coords"], c["u0"], c["B"], 3.0, 1e-3, 1e3)) <= 1e-12
    assert eta.min() == 1e-3 and eta.max() == 1e3
    bu, bp = stokes_rhs_lifted(dims, c["coords"], eta, c["bc"], c["g"], c["f"])
    mg = rift_hierarchy(c, 3, eta, default_options(smoother_precision=32, pre_its=4, post_its=4,
                                                    esteig=(0.0, 0.1, 0.0, 1.1)))
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, bp, rtol=1e-8, maxit=300, restart=100)
    mg.destroy()
    bn = np.linalg.norm(np.concatenate([bu, bp]))
    assert rn <= 1e-8 * bn, its
    m = c["bc"].astype(bool)
    # the identity rows carry the boundary values up to the solver tolerance
    assert np.linalg.norm(u[m] - c["g"][m]) <= 1e-8 * bn
    # masked operator residual
    yu, yp = _oracle_stokes(dims, c["coords"], eta, c["bc"], u, p)
    assert np.linalg.norm(np.concatenate([bu - yu, bp - yp])) <= 2e-8 * bn
    # unconstrained equations: A11 u + A12 p = f at the free dofs, A21 u = 0
    z = np.zeros_like(c["bc"])
    ru = O.a11_tensor(*dims, c["coords"], eta, u) + O.a12_matmult(*dims, c["coords"], z, p) - c["f"]
    rp = O.a21_matmult(*dims, c["coords"], z, u)
    assert np.linalg.norm(ru[~m]) <= 2e-8 * bn
    assert np.linalg.norm(rp) <= 2e-8 * bn


def test_bdb_schur_blocks_match_oracle(gpu):
    """B^T diag(A11)^-1 B Schur blocks: one FGMRES iteration on [0 | r_p]
    returns x_p = alpha Minv r_p (the preconditioner's first direction), so
    x_p must be parallel to the oracle's blocks applied to r_p."""
    m = (6, 4, 2)
    coords, eta, bc, rng = _case(m)
    mg = _hier(m, 2, coords, eta)
    mg.set_schur("bdb")
    E = m[0] * m[1] * m[2]
    rp = rng.uniform(-1, 1, 4 * E)
    bu = np.zeros(3 * F.nnodes(*m))
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, rp, rtol=1e-30, maxit=1, restart=5)
    want = np.einsum("eij,ej->ei", O.schur_bdb_blocks(*m, coords, eta, bc), rp.reshape(E, 4)).reshape(-1)
    alpha = (p @ want) / (want @ want)
    assert np.abs(p).max() > 0 and alpha != 0.0
    assert np.abs(p - alpha * want).max() <= 1e-11 * np.abs(p).max()
    # switching back restores the scaled-mass blocks
    mg.set_schur("mass")
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, rp, rtol=1e-30, maxit=1, restart=5)
    Minv = np.linalg.inv(O.schur_mass_blocks(*m, coords, eta))
    want = np.einsum("eij,ej->ei", Minv, rp.reshape(E, 4)).reshape(-1)
    alpha = (p @ want) / (want @ want)
    assert np.abs(p - alpha * want).max() <= 1e-11 * np.abs(p).max()
    mg.destroy()


def test_rift_bdb_schur_converges(gpu):
    """Rift workload at 16x16x8 (contrast 1e6): with the B^T D^-1 B blocks
    the saddle-point FGMRES reaches 1e-8 in a few tens of iterations (the
    scaled viscous mass needs over a hundred), checked through the oracle's
    masked residual."""
    from ptatin3d_amd.multigrid import powerlaw_viscosity, rift_hierarchy, stokes_rhs_lifted

    dims = (16, 16, 8)
    c = F.config5(*dims)
    eta = powerlaw_viscosity(dims, c["coords"], c["u0"], c["B"], c["nexp"], c["eta_min"], c["eta_max"])
    bu, bp = stokes_rhs_lifted(dims, c["coords"], eta, c["bc"], c["g"], c["f"])
    mg = rift_hierarchy(c, 3, eta, default_options(smoother_precision=32, pre_its=4, post_its=4,
                                                    esteig=(0.0, 0.1, 0.0, 1.1)), schur="bdb")
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, bp, rtol=1e-8, maxit=300, restart=100)
    mg.destroy()
    bn = np.linalg.norm(np.concatenate([bu, bp]))
    assert rn <= 1e-8 * bn and its <= 60, its
    yu, yp = _oracle_stokes(dims, c["coords"], eta, c["bc"], u, p)
    assert np.linalg.norm(np.concatenate([bu - yu, bp - yp])) <= 2e-8 * bn


def test_bdb_schur_singular_block_is_an_error(gpu):
    """2x2x2 box with every velocity dof Dirichlet except the x component of
    the centre node: each element has one free velocity row, so its
    B^T D^-1 B block (rank <= 1) is singular.  set_schur reports it and
    leaves the viscous-mass blocks in place."""
    from ptatin3d_amd._lib import PtmfError

    m = (2, 2, 2)
    coords, eta, _, rng = _case(m)
    bc = np.ones(3 * F.nnodes(*m), np.uint8)
    bc[3 * (2 + 5 * (2 + 5 * 2))] = 0
    mg = GMG(*m, 2)
    mg.set_coordinates(coords)
    mg.set_viscosity(1, eta)
    mg.coarsen_viscosity()
    mg.set_dirichlet(1, bc)
    mg.set_dirichlet(0, F.dirichlet_all_faces(1, 1, 1))
    mg.setup(default_options(smoother_precision=64))
    mg.stokes_setup()
    with pytest.raises(PtmfError, match="singular"):
        mg.set_schur("bdb")
    E = 8
    rp = rng.uniform(-1, 1, 4 * E)
    bu = np.zeros(3 * F.nnodes(*m))
    u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, rp, rtol=1e-30, maxit=1, restart=5)
    want = np.einsum("eij,ej->ei", np.linalg.inv(O.schur_mass_blocks(*m, coords, eta)),
                     rp.reshape(E, 4)).reshape(-1)
    alpha = (p @ want) / (want @ want)
    assert np.abs(p).max() > 0 and np.abs(p - alpha * want).max() <= 1e-11 * np.abs(p).max()
    mg.destroy()
