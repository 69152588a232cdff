"""Assembled level operators and rigid-body modes through the C-ABI
(ptmf_mg_level_assemble, ptmf_mg_level_rigid_body_modes).

The reference assembles a level from its element matrices
(MatAssemble_StokesA_AUU, stokes_assembly.c:115-297): Dirichlet rows and
columns are skipped and a unit diagonal inserted (BCListInsertScaling,
dmda_bcs.c:1035-1060).  The oracle's element matrices
(orc_a11_element_matrix, pinned by a brute-force numpy element matrix in
test_oracle.py) assembled the same way are the expected matrix.  The
near null space the reference attaches (ptatin_driver_linear_ts.c:663-666)
comes from PETSc's MatNullSpaceCreateRigidBody, which is not in the
reference tree: the modes are checked against their closed form and
against A r = 0, not against PETSc's vectors (whose scaling is therefore
parity unpinned; their span is the same)."""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray
from ptatin3d_amd.multigrid import GMG, default_options

pytestmark = pytest.mark.gpu


def _dense_csr(rowptr, colidx, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(rowptr[r], rowptr[r + 1]):
            A[r, colidx[k]] += vals[k]
    return A


def _oracle_assembled(m, coords, eta, bc):
    """Element matrices of the oracle summed into a dense matrix, Dirichlet
    rows and columns removed and a unit diagonal inserted."""
    mx, my, mz = m
    X, Y = 2 * mx + 1, 2 * my + 1
    n = 3 * F.nnodes(*m)
    A = np.zeros((n, n))
    for e in range(mx * my * mz):
        ei, ej, ek = e % mx, (e // mx) % my, e // (mx * my)
        nid = [(2 * ei + i) + X * ((2 * ej + j) + Y * (2 * ek + k))
               for k in range(3) for j in range(3) for i in range(3)]
        dofs = np.array([3 * g + d for g in nid for d in range(3)])
        A[np.ix_(dofs, dofs)] += O.a11_element_matrix(mx, my, mz, coords, eta, e)
    d = np.asarray(bc, dtype=bool)
    A[d, :] = 0.0
    A[:, d] = 0.0
    A[d, d] = 1.0
    return A


def _check_structure(rowptr, colidx, vals, bc):
    n = rowptr.size - 1
    d = np.asarray(bc, dtype=bool)
    assert rowptr[0] == 0 and np.all(np.diff(rowptr) >= 0) and rowptr[-1] == colidx.size
    for r in range(n):
        cols = colidx[rowptr[r]:rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0), r  # ascending, no duplicates
        if d[r]:
            assert cols.tolist() == [r] and vals[rowptr[r]] == 1.0
        else:
            assert not np.any(d[cols]), r  # no Dirichlet columns in free rows


@pytest.mark.parametrize("m,bckind", [((3, 4, 2), "faces"), ((4, 3, 2), "rift"), ((2, 2, 3), "faces")])
def test_level_assemble_matches_oracle_element_assembly(gpu, m, bckind):
    rng = np.random.default_rng(sum(m) + len(bckind))
    coords = F.perturb_interior(F.box_coords(*m, hi=(2.0, 1.5, 1.0)), *m, 0.1, rng)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-2, 1e2, rng)
    bc = F.dirichlet_all_faces(*m) if bckind == "faces" else F.rift_bc(*m)[0]
    print(f"assemble: box {m} bc {bckind}", flush=True)
    mg = GMG(*m, 1)
    mg.set_coordinates(coords)
    mg.set_viscosity(0, eta)
    mg.set_dirichlet(0, bc)
    mg.setup(default_options(smoother_precision=64))
    rowptr, colidx, vals = mg.level_assemble(0)
    n = 3 * F.nnodes(*m)
    _check_structure(rowptr, colidx, vals, bc)
    A = _dense_csr(rowptr, colidx, vals, n)
    Ao = _oracle_assembled(m, coords, eta, bc)
    assert np.max(np.abs(A - Ao)) <= 1e-12 * np.max(np.abs(Ao))
    # every structural entry of the oracle matrix is present (nothing dropped)
    assert np.count_nonzero(Ao[A == 0.0]) == 0
    # the matrix is the MatMult's operator
    x = rng.uniform(-1, 1, n)
    y = O.a11_matmult(*m, coords, eta, bc, x)
    assert np.max(np.abs(A @ x - y)) <= 1e-12 * np.max(np.abs(y))
    mg.destroy()


def test_level_assemble_galerkin_and_rediscretised_levels(gpu):
    """Two levels, the coarse one a Galerkin stencil: each level's CSR matrix
    applies as the level's own operator, and is symmetric."""
    m = (4, 4, 4)
    rng = np.random.default_rng(11)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    eta = F.eta_loguniform(64, 1e-2, 1e2, rng)
    mg = GMG(*m, 2)
    mg.set_coordinates(coords)
    mg.set_viscosity(1, eta)
    mg.coarsen_viscosity()
    mg.set_dirichlet_all_faces()
    mg.setup(default_options(smoother_precision=64, galerkin_levels=1))
    for lvl in (0, 1):
        n = mg.level_n(lvl)
        rowptr, colidx, vals = mg.level_assemble(lvl)
        A = _dense_csr(rowptr, colidx, vals, n)
        assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
        x = rng.uniform(-1, 1, n)
        dx, dy = DeviceArray.from_host(x), DeviceArray(n)
        mg.level_apply(lvl, dx, dy)
        y = dy.download()
        dx.free()
        dy.free()
        assert np.max(np.abs(A @ x - y)) <= 1e-12 * np.max(np.abs(y)), lvl
    mg.destroy()


def test_rigid_body_modes(gpu):
    """Closed form from the level's coordinates, and in the null space of a
    level without Dirichlet rows (A r = 0 up to rounding)."""
    m = (4, 4, 2)
    rng = np.random.default_rng(3)
    coords = F.perturb_interior(F.box_coords(*m, lo=(1.0, -2.0, 0.5), hi=(3.0, 0.0, 1.5)), *m, 0.1, rng)
    eta = F.eta_loguniform(32, 1e-2, 1e2, rng)
    mg = GMG(*m, 2)
    mg.set_coordinates(coords)
    mg.set_viscosity(1, eta)
    mg.coarsen_viscosity()
    # Dirichlet rows on the coarse level only (its factor needs them); the
    # fine level keeps the unconstrained operator
    mc = (2, 2, 1)
    n = 3 * F.nnodes(*m)
    mg.set_dirichlet(0, F.dirichlet_all_faces(*mc))
    mg.set_dirichlet(1, np.zeros(n, dtype=np.uint8))
    mg.setup(default_options(smoother_precision=64))
    R = mg.level_rigid_body_modes(1)
    assert R.shape == (6, n)
    x, y, z = coords[0::3], coords[1::3], coords[2::3]
    E = np.zeros((6, n))
    for c in range(3):
        E[c, c::3] = 1.0
    E[3, 0::3], E[3, 1::3] = -y, x
    E[4, 1::3], E[4, 2::3] = -z, y
    E[5, 0::3], E[5, 2::3] = z, -x
    assert np.array_equal(R, E)
    rowptr, colidx, vals = mg.level_assemble(1)
    A = _dense_csr(rowptr, colidx, vals, n)
    scale = np.max(np.abs(A))
    for k in range(6):
        dx, dy = DeviceArray.from_host(R[k]), DeviceArray(n)
        mg.level_apply(1, dx, dy)
        r = dy.download()
        dx.free()
        dy.free()
        assert np.max(np.abs(r)) <= 1e-11 * scale * np.max(np.abs(R[k])), k
        assert np.max(np.abs(A @ R[k])) <= 1e-11 * scale * np.max(np.abs(R[k])), k
    mg.destroy()
