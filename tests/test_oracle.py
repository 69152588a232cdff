"""Pins the CPU oracle (oracle/ptmf_oracle.c) before it is trusted as the checker.

The reference cannot be built here (it needs PETSc), so each oracle function is
held to checks that a plausible mistake (dropped term, wrong sign or index,
transposed operand) would fail:

  a11_tensor   == a11_ref (independent per-qp dense restatement),
               == assembled element matrices (the reference's own MF-vs-ASM
                  invariant, test_stokes_operators.c:225-319, tol 1e-13),
               symmetric, kills the 6 rigid-body modes on curved meshes,
               uT A u = 2 eta |E|^2 Vol for linear u on affine meshes;
  element matrix == independent numpy brute force of int 2 eta eps(u):eps(v);
  matmult / asm_matmult: two independent codings of the Dirichlet semantics;
  diagonal     == diag of the summed element matrices, unit Dirichlet slots;
  a12 / a21    adjoint to each other, == numpy brute force, closed forms;
  transfers    reproduce constants and lattice-linear fields, R == P^T;
  chebyshev    == textbook Chebyshev semi-iteration (same polynomial) and
               obeys the Chebyshev error bound with exact eigenvalue bounds;
  powerlaw_eta affine u on a curved mesh has grad u = G exactly (closed-form
               eII), u = (x^2, y z, 0) on a box gives eII at the quadrature
               points' own positions (pins qp order and element order), the
               1e-17 strain-rate floor and the clipping bounds;
  q1 hierarchy == numpy restatement with explicit weight matrices (the
               projection weights, R = P^T from the pinned restrict_add, the
               clamped Q1 interpolation), constants reproduced exactly,
               values stay inside the fine range.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F

from conftest import rel_inf

G3 = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
W3 = np.array([5.0, 8.0, 5.0]) / 9.0


def _n1(t):
    return np.array([0.5 * t * (t - 1.0), (1.0 + t) * (1.0 - t), 0.5 * t * (t + 1.0)])


def _dn1(t):
    return np.array([t - 0.5, -2.0 * t, t + 0.5])


def brute_element(xe, eta_e):
    """81x81 stiffness of int 2 eta eps(u):eps(v) by explicit strain tensors."""
    K = np.zeros((81, 81))
    for c in range(3):
        for b in range(3):
            for a in range(3):
                q = a + 3 * b + 9 * c
                xi = (G3[a], G3[b], G3[c])
                n = [_n1(t) for t in xi]
                dn = [_dn1(t) for t in xi]
                gref = np.zeros((3, 27))
                for k in range(3):
                    for j in range(3):
                        for i in range(3):
                            m = i + 3 * j + 9 * k
                            gref[0, m] = dn[0][i] * n[1][j] * n[2][k]
                            gref[1, m] = n[0][i] * dn[1][j] * n[2][k]
                            gref[2, m] = n[0][i] * n[1][j] * dn[2][k]
                J = gref @ xe  # J[a,d] = dx_d/dxi_a
                gx = np.linalg.solve(J, gref)  # physical gradients [3,27]
                detJ = np.linalg.det(J)
                eps = np.zeros((81, 3, 3))
                for m in range(27):
                    for d in range(3):
                        E = np.zeros((3, 3))
                        E[:, d] += 0.5 * gx[:, m]
                        E[d, :] += 0.5 * gx[:, m]
                        eps[3 * m + d] = E
                K += W3[a] * W3[b] * W3[c] * detJ * 2.0 * eta_e[q] * np.einsum("ikl,jkl->ij", eps, eps)
    return K, None


@pytest.fixture(scope="module")
def curved():
    m = (3, 2, 4)
    rng = np.random.default_rng(7)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.15, rng)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-2, 1e2, rng)
    bc = F.dirichlet_all_faces(*m)
    return m, coords, eta, bc, rng


def test_tensor_equals_ref_and_assembled(curved):
    m, coords, eta, bc, rng = curved
    u = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    yt = O.a11_tensor(*m, coords, eta, u)
    yr = O.a11_ref(*m, coords, eta, u)
    assert rel_inf(yt, yr) < 1e-13
    nobc = np.zeros_like(bc)
    ya = O.a11_asm_matmult(*m, coords, eta, nobc, u)
    assert rel_inf(yt, ya) < 1e-13


def test_element_matrix_brute_force(curved):
    m, coords, eta, bc, rng = curved
    nid = F.element_node_ids(*m)
    for e in (0, 5, m[0] * m[1] * m[2] - 1):
        Ae = O.a11_element_matrix(*m, coords, eta, e)
        K, _ = brute_element(coords.reshape(-1, 3)[nid[e]], eta[27 * e: 27 * e + 27])
        assert rel_inf(Ae, K) < 1e-12
        assert np.allclose(Ae, Ae.T, rtol=0, atol=1e-14 * np.abs(Ae).max())
        ev = np.linalg.eigvalsh(Ae)
        # exactly six rigid-body modes in the kernel, the rest positive
        assert (np.abs(ev[:6]) < 1e-10 * ev[-1]).all() and ev[6] > 1e-8 * ev[-1]


def test_symmetry_and_rigid_modes(curved):
    m, coords, eta, bc, rng = curved
    n = 3 * F.nnodes(*m)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    Ax = O.a11_tensor(*m, coords, eta, x)
    Ay = O.a11_tensor(*m, coords, eta, y)
    assert abs(y @ Ax - x @ Ay) < 1e-12 * abs(y @ Ax)
    c = coords.reshape(-1, 3)
    scale = np.abs(Ax).max()
    modes = [np.tile([1.0, 0, 0], len(c)), np.tile([0, 1.0, 0], len(c)), np.tile([0, 0, 1.0], len(c))]
    for (i, j) in ((0, 1), (0, 2), (1, 2)):
        r = np.zeros_like(c)
        r[:, i] = -c[:, j]
        r[:, j] = c[:, i]
        modes.append(r.reshape(-1))
    for r in modes:
        assert np.abs(O.a11_tensor(*m, coords, eta, r)).max() < 1e-12 * scale * np.abs(r).max()


def test_linear_field_energy_closed_form():
    m = (2, 3, 2)
    coords = F.box_coords(*m, hi=(1.5, 0.7, 2.0))
    eta0 = 3.7
    eta = np.full(27 * m[0] * m[1] * m[2], eta0)
    E = np.array([[0.3, 0.1, -0.2], [0.1, -0.5, 0.4], [-0.2, 0.4, 0.9]])
    c = coords.reshape(-1, 3)
    u = (c @ E.T).reshape(-1)
    Au = O.a11_tensor(*m, coords, eta, u)
    vol = 1.5 * 0.7 * 2.0
    assert abs(u @ Au - 2 * eta0 * np.sum(E * E) * vol) < 1e-12 * abs(u @ Au)


def test_matmult_bc_semantics(curved):
    m, coords, eta, bc, rng = curved
    x = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    y1 = O.a11_matmult(*m, coords, eta, bc, x)
    y2 = O.a11_asm_matmult(*m, coords, eta, bc, x)
    assert rel_inf(y1, y2) < 1e-13
    d = bc.astype(bool)
    assert np.array_equal(y1[d], x[d])
    # columns of Dirichlet dofs are removed: changing x there only changes
    # those rows (up to the scatter order of the OpenMP critical section)
    x2 = x.copy()
    x2[d] += 1.0
    y3 = O.a11_matmult(*m, coords, eta, bc, x2)
    assert rel_inf(y3[~d], y1[~d]) < 1e-14


def test_reference_test_vector(curved):
    """The reference's operator-test input f = x+y+z+comp+3 through both paths."""
    m, coords, eta, bc, rng = curved
    f = F.reference_test_vector(coords)
    assert rel_inf(O.a11_matmult(*m, coords, eta, bc, f), O.a11_asm_matmult(*m, coords, eta, bc, f)) < 1e-13


def test_diagonal(curved):
    m, coords, eta, bc, rng = curved
    nid = F.element_node_ids(*m)
    d_ref = np.zeros(3 * F.nnodes(*m))
    for e in range(m[0] * m[1] * m[2]):
        Ae = O.a11_element_matrix(*m, coords, eta, e)
        for n in range(27):
            for k in range(3):
                d_ref[3 * nid[e, n] + k] += Ae[3 * n + k, 3 * n + k]
    d_ref[bc.astype(bool)] = 1.0
    d = O.a11_diagonal(*m, coords, eta, bc)
    assert rel_inf(d, d_ref) < 1e-13
    # probe a few free dofs against e_i^T A e_i of the tensor operator
    free = np.flatnonzero(bc == 0)
    for i in free[:: max(1, len(free) // 5)]:
        ei = np.zeros(3 * F.nnodes(*m))
        ei[i] = 1.0
        assert abs(O.a11_matmult(*m, coords, eta, bc, ei)[i] - d[i]) < 1e-12 * abs(d[i])


def _brute_b(m, coords):
    """dense B (3N x 4E) of -int q div v with the P1disc basis, numpy"""
    nid = F.element_node_ids(*m)
    N = F.nnodes(*m)
    E = m[0] * m[1] * m[2]
    Bm = np.zeros((3 * N, 4 * E))
    X = coords.reshape(-1, 3)
    for e in range(E):
        xe = X[nid[e]]
        avg = xe.mean(axis=0)
        L = np.array([xe[14, 0] - xe[12, 0], xe[16, 1] - xe[10, 1], xe[22, 2] - xe[4, 2]])
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    xi = (G3[a], G3[b], G3[c])
                    n = [_n1(t) for t in xi]
                    dn = [_dn1(t) for t in xi]
                    Nv = np.einsum("k,j,i->kji", n[2], n[1], n[0]).reshape(27)
                    gref = np.stack([
                        np.einsum("k,j,i->kji", n[2], n[1], dn[0]).reshape(27),
                        np.einsum("k,j,i->kji", n[2], dn[1], n[0]).reshape(27),
                        np.einsum("k,j,i->kji", dn[2], n[1], n[0]).reshape(27)])
                    J = gref @ xe
                    gx = np.linalg.solve(J, gref)
                    w = W3[a] * W3[b] * W3[c] * np.linalg.det(J)
                    xq = Nv @ xe
                    Np = np.concatenate([[1.0], (xq - avg) / L])
                    for mm in range(27):
                        for d in range(3):
                            Bm[3 * nid[e, mm] + d, 4 * e:4 * e + 4] += -w * gx[d, mm] * Np
    return Bm


def test_a12_a21(curved):
    m, coords, eta, bc, rng = curved
    E = m[0] * m[1] * m[2]
    N = F.nnodes(*m)
    Bm = _brute_b(m, coords)
    p = rng.uniform(-1, 1, 4 * E)
    u = rng.uniform(-1, 1, 3 * N)
    d = bc.astype(bool)
    y12 = O.a12_matmult(*m, coords, bc, p)
    r12 = Bm @ p
    r12[d] = 0.0
    assert rel_inf(y12, r12) < 1e-12
    um = u.copy()
    um[d] = 0.0
    assert rel_inf(O.a21_matmult(*m, coords, bc, u), Bm.T @ um) < 1e-12
    # adjoint on the BC-masked spaces
    assert abs(p @ O.a21_matmult(*m, coords, bc, u) - um @ O.a12_matmult(*m, coords, bc, p)) < 1e-12 * np.abs(y12).max() * np.abs(u).sum()


def test_a21_closed_form_affine():
    """u = x (div u = 3) on an affine box: Y_p[e] = (-3 vol_e, 0, 0, 0)."""
    m = (2, 2, 3)
    coords = F.box_coords(*m, hi=(1.0, 2.0, 1.5))
    bc = np.zeros(3 * F.nnodes(*m), dtype=np.uint8)
    yp = O.a21_matmult(*m, coords, bc, coords).reshape(-1, 4)
    vol = (1.0 / 2) * (2.0 / 2) * (1.5 / 3)
    assert np.allclose(yp[:, 0], -3.0 * vol, rtol=1e-13)
    assert np.abs(yp[:, 1:]).max() < 1e-13


def test_transfers():
    mc = (2, 3, 2)
    mf = tuple(2 * v for v in mc)
    rng = np.random.default_rng(3)
    cc = F.box_coords(*mc)
    cf = F.box_coords(*mf)
    # constants and lattice-linear fields are reproduced exactly
    lin = lambda c: (c.reshape(-1, 3) @ np.array([[1.0, 2.0, -1.0], [0.5, 0.0, 3.0], [2.0, -1.0, 1.0]])).reshape(-1)
    yf = O.prolong_add(*mc, lin(cc), np.zeros(3 * F.nnodes(*mf)))
    assert np.abs(yf - lin(cf)).max() < 1e-13
    # restriction is the transpose
    xf = rng.uniform(-1, 1, 3 * F.nnodes(*mf))
    yc = rng.uniform(-1, 1, 3 * F.nnodes(*mc))
    Rxf = O.restrict_add(*mc, xf, np.zeros(3 * F.nnodes(*mc)))
    Pyc = O.prolong_add(*mc, yc, np.zeros(3 * F.nnodes(*mf)))
    assert abs(Rxf @ yc - xf @ Pyc) < 1e-12 * abs(xf @ Pyc)
    # row sums of R: interior coarse nodes receive weight (1 + 2*1/2)^3 = 8
    ones = O.restrict_add(*mc, np.ones(3 * F.nnodes(*mf)), np.zeros(3 * F.nnodes(*mc)))
    X, Y, Z = 2 * mc[0] + 1, 2 * mc[1] + 1, 2 * mc[2] + 1
    interior = (1 + X * (1 + Y * 1))
    assert ones[3 * interior] == 8.0
    # coordinate injection
    assert np.array_equal(O.inject_coords(*mc, cf), cc)


def _dense(m, coords, eta, bc):
    n = 3 * F.nnodes(*m)
    A = np.zeros((n, n))
    for i in range(n):
        ei = np.zeros(n)
        ei[i] = 1.0
        A[:, i] = O.a11_matmult(*m, coords, eta, bc, ei)
    return A


def test_chebyshev_matches_textbook_and_bound():
    m = (2, 2, 2)
    rng = np.random.default_rng(11)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    eta = F.eta_loguniform(8, 1e-1, 1e1, rng)
    bc = F.dirichlet_all_faces(*m)
    A = _dense(m, coords, eta, bc)
    d = O.a11_diagonal(*m, coords, eta, bc)
    assert np.allclose(np.diag(A), d, rtol=1e-12)
    Dm = 1.0 / np.sqrt(d)
    lam = np.linalg.eigvalsh(Dm[:, None] * A * Dm[None, :])
    emin, emax = lam[0], lam[-1]
    b = rng.uniform(-1, 1, A.shape[0])
    x0 = rng.uniform(-1, 1, A.shape[0])
    its = 6
    x = O.cheb_jacobi(*m, coords, eta, bc, d, emin, emax, its, b, x0)
    # textbook Chebyshev semi-iteration (Saad, Iterative Methods, Alg. 12.1)
    theta, delta = 0.5 * (emax + emin), 0.5 * (emax - emin)
    sigma = theta / delta
    rho = 1.0 / sigma
    xs = x0.copy()
    r = b - A @ xs
    dd = (r / d) / theta
    for _ in range(its + 1):
        xs = xs + dd
        r = r - A @ dd
        rho_n = 1.0 / (2 * sigma - rho)
        dd = rho_n * rho * dd + (2 * rho_n / delta) * (r / d)
        rho = rho_n
    assert rel_inf(x, xs) < 1e-11
    # error bound in the A-norm
    xstar = np.linalg.solve(A, b)
    en = lambda v: np.sqrt((v - xstar) @ A @ (v - xstar))
    kappa = emax / emin
    q = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    k = its + 1
    assert en(x) <= 2 * q ** k / (1 + q ** (2 * k)) * en(x0) * (1 + 1e-9)


def _powerlaw(B, sr, n, lo, hi):
    return np.clip(B * np.maximum(sr, 1e-17) ** (1.0 / n - 1.0), lo, hi)


def test_powerlaw_affine_curved(curved):
    m, coords, eta, bc, rng = curved
    G = rng.uniform(-1, 1, (3, 3))
    u = (coords.reshape(-1, 3) @ G.T).reshape(-1)
    D = 0.5 * (G + G.T)
    sr = np.sqrt(0.5 * np.sum(D * D))
    B = rng.uniform(1, 10, 27 * m[0] * m[1] * m[2])
    got = O.powerlaw_eta(*m, coords, u, B, 3.0, 1e-6, 1e6)
    assert rel_inf(got, _powerlaw(B, sr, 3.0, 1e-6, 1e6)) <= 1e-12


def test_powerlaw_quadratic_at_qp_positions():
    m = (3, 2, 4)
    coords = F.box_coords(*m, hi=(1.3, 0.8, 1.1))
    c = coords.reshape(-1, 3)
    u = np.stack([c[:, 0] ** 2, c[:, 1] * c[:, 2], np.zeros(len(c))], axis=1).reshape(-1)
    qx = F.qp_coordinates(coords, *m).reshape(-1, 3)  # element-major [27e+q]
    x, y, z = qx[:, 0], qx[:, 1], qx[:, 2]
    # D = [[2x,0,0],[0,z,y/2],[0,y/2,0]]
    sr = np.sqrt(0.5 * (4 * x * x + z * z + 2 * (0.5 * y) ** 2))
    B = np.full(qx.shape[0], 2.5)
    got = O.powerlaw_eta(*m, coords, u, B, 2.0, 1e-9, 1e9)
    assert rel_inf(got, _powerlaw(B, sr, 2.0, 1e-9, 1e9)) <= 1e-12
    # Newtonian limit n = 1 returns B; zero velocity hits the floor and then eta_max
    assert np.array_equal(O.powerlaw_eta(*m, coords, u, B, 1.0, 1e-9, 1e9), B)
    z0 = O.powerlaw_eta(*m, coords, np.zeros_like(u), B, 3.0, 1e-3, 1e3)
    assert np.all(z0 == 1e3)
    z1 = O.powerlaw_eta(*m, coords, np.zeros_like(u), B, 3.0, 1e-3, 1e40)
    assert rel_inf(z1, B * 1e-17 ** (-2.0 / 3.0)) <= 1e-14


def _q1_hierarchy_numpy(dims, eta, inject=False):
    """Independent restatement: explicit per-element weight tables."""
    g = F.GAUSS
    f1 = np.stack([0.5 * (1 - g), 0.5 * (1 + g)], axis=1)  # [qp 1-D, corner a]
    wq1 = np.einsum("ia,jb,kc->kjicba", f1, f1, f1).reshape(27, 8)  # [q, corner]
    corners = np.array([0, 2, 6, 8, 18, 20, 24, 26])  # corner (a,b,c) -> local node 2a+6b+18c
    Wt = np.ones((27, 27))  # [q, local node]
    Wt[:, corners] = wq1
    mx, my, mz = dims[-1]
    nid = F.element_node_ids(mx, my, mz)
    e27 = eta.reshape(-1, 27)
    A = np.zeros(F.nnodes(mx, my, mz))
    Bw = np.zeros_like(A)
    np.add.at(A, nid, e27 @ Wt)
    np.add.at(Bw, nid, np.broadcast_to(Wt.sum(axis=0), nid.shape))
    nodal = A / Bw
    out = [None] * len(dims)
    out[-1] = eta
    for l in range(len(dims) - 1, 0, -1):
        cm = dims[l - 1]
        if inject:
            X, Y, Z = (2 * dims[l][0] + 1, 2 * dims[l][1] + 1, 2 * dims[l][2] + 1)
            nodal = nodal.reshape(Z, Y, X)[::2, ::2, ::2].reshape(-1)
        else:
            v3 = np.stack([nodal, np.ones_like(nodal), np.zeros_like(nodal)], axis=1).reshape(-1)
            r = O.restrict_add(*cm, v3, np.zeros(3 * F.nnodes(*cm))).reshape(-1, 3)
            nodal = r[:, 0] / r[:, 1]
        nidc = F.element_node_ids(*cm)
        ev = nodal[nidc]  # [E, 27]
        q = ev[:, corners] @ wq1.T  # [E, 27]
        q = np.clip(q, ev.min(axis=1, keepdims=True), ev.max(axis=1, keepdims=True))
        out[l - 1] = q.reshape(-1)
    return out


@pytest.mark.parametrize("inject", [False, True])
def test_q1_hierarchy_against_numpy(inject):
    dims = [(2, 1, 3), (4, 2, 6), (8, 4, 12)]
    rng = np.random.default_rng(3)
    eta = F.eta_loguniform(8 * 4 * 12, 1e-3, 1e3, rng)
    got = O.q1_projection_hierarchy(dims, eta, inject)
    ref = _q1_hierarchy_numpy(dims, eta, inject)
    for g, r in zip(got, ref):
        assert rel_inf(g, r) <= 1e-13
        assert g.min() >= eta.min() * (1 - 1e-14) and g.max() <= eta.max() * (1 + 1e-14)
    const = O.q1_projection_hierarchy(dims, np.full_like(eta, 3.25), inject)
    for g in const:
        assert np.abs(g - 3.25).max() <= 1e-14


def _golden_cases():
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "viscous_sinker_operators.json")
    return json.load(open(path))["cases"]


def _close9(v, ref):
    """Equal to the 9 significant digits the reference prints (within half a
    unit of the last printed digit)."""
    return abs(v - ref) <= 0.5 * 10.0 ** (np.floor(np.log10(abs(ref))) - 8)


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: "x".join(map(str, c["mesh"])) + f"_eta{c['eta1']:g}")
def test_reference_golden_values_viscous_sinker(case):
    """The reference's own operator test (viscous_sinker on the Gaussian-bump
    mesh, 2x2x2 lattice markers projected to the quadrature points, free slip
    with a free top): y.y of A11 x, A21 x, A12 p and diag(A11) equal the
    printed expected values (test_stokes_operators.c:225-580)."""
    m = tuple(case["mesh"])
    coords = F.gaussian_bump_coords(*m)
    elem, xi, val = F.sinker_markers(coords, *m, eta1=case["eta1"])
    eta = O.q1_to_qp(*m, O.q1_nodal_from_points(*m, elem, xi, val))
    bc = F.sinker_bc(*m)
    xu = F.reference_test_vector(coords)
    y = O.a11_matmult(*m, coords, eta, bc, xu)
    assert _close9(y @ y, case["A11"]), y @ y
    y = O.a21_matmult(*m, coords, bc, xu)
    assert _close9(y @ y, case["A21"]), y @ y
    if case["A12"] is not None:
        y = O.a12_matmult(*m, coords, bc, F.reference_pressure_test_vector(*m))
        assert _close9(y @ y, case["A12"]), y @ y
    d = O.a11_diagonal(*m, coords, eta, bc)
    assert _close9(d @ d, case["diagA11"]), d @ d


def test_point_projection_at_quadrature_points_is_qp_projection():
    m = (3, 2, 4)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-2, 1e2, np.random.default_rng(2))
    qx = np.tile(np.stack(np.meshgrid(F.GAUSS, F.GAUSS, F.GAUSS, indexing="ij")[::-1], -1).reshape(27, 3),
                 (m[0] * m[1] * m[2], 1))
    el = np.repeat(np.arange(m[0] * m[1] * m[2]), 27)
    assert rel_inf(O.q1_nodal_from_points(*m, el, qx, eta), O.q1_nodal_from_qp(*m, eta)) <= 1e-14


def _brute_schur_mass(m, coords, eta):
    """element blocks -sum_q w detJ (1/eta_q) NIp_i NIp_j, numpy (P1GRel basis
    at the physical quadrature point, geometry as in _brute_b)"""
    nid = F.element_node_ids(*m)
    E = m[0] * m[1] * m[2]
    X = coords.reshape(-1, 3)
    out = np.zeros((E, 4, 4))
    for e in range(E):
        xe = X[nid[e]]
        avg = xe.mean(axis=0)
        L = np.array([xe[14, 0] - xe[12, 0], xe[16, 1] - xe[10, 1], xe[22, 2] - xe[4, 2]])
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    xi = (G3[a], G3[b], G3[c])
                    n = [_n1(t) for t in xi]
                    dn = [_dn1(t) for t in xi]
                    Nv = np.einsum("k,j,i->kji", n[2], n[1], n[0]).reshape(27)
                    gref = np.stack([
                        np.einsum("k,j,i->kji", n[2], n[1], dn[0]).reshape(27),
                        np.einsum("k,j,i->kji", n[2], dn[1], n[0]).reshape(27),
                        np.einsum("k,j,i->kji", dn[2], n[1], n[0]).reshape(27)])
                    w = W3[a] * W3[b] * W3[c] * np.linalg.det(gref @ xe)
                    Np = np.concatenate([[1.0], (Nv @ xe - avg) / L])
                    out[e] -= (w / eta[27 * e + a + 3 * b + 9 * c]) * np.outer(Np, Np)
    return out


def test_schur_mass_blocks_brute_force(curved):
    """orc_schur_mass_blocks (stokes_assembly.c:386-409) against an independent
    numpy quadrature of -int (1/eta) NIp_i NIp_j on curved elements with
    quadrature-point viscosity; a flipped sign or eta in place of 1/eta fails."""
    m, coords, eta, bc, rng = curved
    Me = O.schur_mass_blocks(*m, coords, eta).reshape(-1, 4, 4)
    ref = _brute_schur_mass(m, coords, eta)
    assert np.abs(Me - ref).max() <= 1e-13 * np.abs(ref).max()
    # negative definite, and the planted mistakes are caught
    assert all(np.linalg.eigvalsh(-b).min() > 0 for b in Me)
    assert np.abs(-Me - ref).max() > 1e-3 * np.abs(ref).max()
    assert np.abs(_brute_schur_mass(m, coords, 1.0 / eta) - Me).max() > 1e-3 * np.abs(ref).max()


def test_schur_mass_blocks_closed_form_affine():
    """Axis-aligned bricks hx x hy x hz with constant eta: NIp = {1, (x-xc)/hx,
    (y-yc)/hy, (z-zc)/hz}, so M_e = -(V/eta) diag(1, 1/12, 1/12, 1/12) exactly
    (3-point Gauss integrates the quadratics)."""
    m = (2, 3, 2)
    coords = F.box_coords(*m, hi=(1.0, 2.0, 1.5))
    E = m[0] * m[1] * m[2]
    eta0 = 7.5
    Me = O.schur_mass_blocks(*m, coords, np.full(27 * E, eta0)).reshape(-1, 4, 4)
    V = (1.0 / 2) * (2.0 / 3) * (1.5 / 2)
    ref = -(V / eta0) * np.diag([1.0, 1.0 / 12, 1.0 / 12, 1.0 / 12])
    for e in range(E):
        assert np.allclose(Me[e], ref, rtol=1e-13, atol=1e-16)


def test_schur_bdb_blocks_brute_force():
    """Element blocks of -B^T D^-1 B against the dense product of the
    assembled A12 (one oracle product per pressure unit vector) and the
    diagonal, Dirichlet rows dropped; every block is the negated SPD inverse."""
    m = (3, 2, 2)
    rng = np.random.default_rng(1)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    E = m[0] * m[1] * m[2]
    eta = F.eta_loguniform(E, 1e-2, 1e2, rng)
    bc = F.dirichlet_all_faces(*m)
    Minv = O.schur_bdb_blocks(*m, coords, eta, bc)
    B = np.stack([O.a12_matmult(*m, coords, bc, np.eye(4 * E)[k]) for k in range(4 * E)], 1)
    d = O.a11_diagonal(*m, coords, eta, bc)
    fr = ~bc.astype(bool)
    assert np.abs(B[~fr]).max() == 0.0
    S = B[fr].T @ (B[fr] / d[fr, None])
    for e in range(E):
        blk = S[4 * e:4 * e + 4, 4 * e:4 * e + 4]
        assert np.allclose(np.linalg.inv(-blk), Minv[e], rtol=1e-12, atol=0)
        assert np.linalg.eigvalsh(-Minv[e]).min() > 0
    # the assembled B^T D^-1 B couples neighbours; the blocks are its diagonal
    assert np.abs(S[0:4, 4:8]).max() > 0


def test_subbox_locality_of_the_unmasked_products():
    """The property the full-size GPU parity tests rest on: without Dirichlet
    masking, the A11 and A12 sums at the complete nodes of a sub-box (strictly
    inside it, or on a domain face it touches) and the A21 rows of its
    elements depend only on the sub-box's elements."""
    from conftest import complete_nodes, sub_elems, sub_nodes

    m = (9, 8, 7)
    rng = np.random.default_rng(31)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    E = m[0] * m[1] * m[2]
    eta = F.eta_loguniform(E, 1e-3, 1e3, rng)
    u = rng.uniform(-1, 1, coords.size)
    p = rng.uniform(-1, 1, 4 * E)
    zf = np.zeros(coords.size, np.uint8)
    yu = O.a11_tensor(*m, coords, eta, u) + O.a12_matmult(*m, coords, zf, p)
    yp = O.a21_matmult(*m, coords, zf, u)
    for lo, n in (((3, 2, 2), (4, 3, 3)), ((0, 0, 0), (3, 4, 2)), ((6, 5, 4), (3, 3, 3))):
        sc = sub_nodes(m, coords, lo, n).reshape(-1)
        su = sub_nodes(m, u, lo, n).reshape(-1)
        se = sub_elems(m, eta, lo, n, 27).reshape(-1)
        sp = sub_elems(m, p, lo, n, 4).reshape(-1)
        zs = np.zeros(sc.size, np.uint8)
        ys = (O.a11_tensor(*n, sc, se, su) + O.a12_matmult(*n, sc, zs, sp)).reshape(
            2 * n[2] + 1, 2 * n[1] + 1, 2 * n[0] + 1, 3)
        ok = complete_nodes(m, lo, n)
        assert rel_inf(sub_nodes(m, yu, lo, n)[ok], ys[ok]) <= 1e-13
        # the incomplete nodes do differ (the check is not vacuous)
        if not ok.all():
            assert np.abs(sub_nodes(m, yu, lo, n)[~ok] - ys[~ok]).max() > 1e-6
        assert rel_inf(sub_elems(m, yp, lo, n, 4).reshape(-1), O.a21_matmult(*n, sc, zs, su)) <= 1e-13
