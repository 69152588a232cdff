"""Seeded synthetic meshes, viscosity fields and vectors for the Q2 box workloads.

Everything is deterministic in (sizes, seed), so the GPU operator and the CPU
oracle are fed identical arrays.  Layouts follow include/ptmf.h.
"""
from __future__ import annotations

import numpy as np

SEED = 20261202
GAUSS = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])


def nnodes(mx: int, my: int, mz: int) -> int:
    return (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)


def box_coords(mx, my, mz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Affine box, node-interlaced coordinates (3N)."""
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    out = np.empty((Z, Y, X, 3))
    out[..., 0] = np.linspace(lo[0], hi[0], X)[None, None, :]
    out[..., 1] = np.linspace(lo[1], hi[1], Y)[None, :, None]
    out[..., 2] = np.linspace(lo[2], hi[2], Z)[:, None, None]
    return out.reshape(-1)


def boundary_node_flags(mx, my, mz) -> np.ndarray:
    """Boolean per node: lies on any face of the box."""
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    b = np.zeros((Z, Y, X), dtype=bool)
    b[0] = b[-1] = True
    b[:, 0] = b[:, -1] = True
    b[:, :, 0] = b[:, :, -1] = True
    return b.reshape(-1)


def perturb_interior(coords, mx, my, mz, frac=0.1, rng=None) -> np.ndarray:
    """Move every interior node by U(-frac*h, frac*h) per coordinate, h the
    node spacing of each direction."""
    rng = rng or np.random.default_rng(SEED)
    c = coords.reshape(-1, 3).copy()
    h = np.array([1.0 / (2 * mx), 1.0 / (2 * my), 1.0 / (2 * mz)])
    h = h * (c.max(axis=0) - c.min(axis=0))
    inner = ~boundary_node_flags(mx, my, mz)
    c[inner] += rng.uniform(-1.0, 1.0, size=(inner.sum(), 3)) * (frac * h)
    return c.reshape(-1)


def dirichlet_all_faces(mx, my, mz) -> np.ndarray:
    """All three velocity components fixed on all six faces (3N bytes)."""
    b = boundary_node_flags(mx, my, mz)
    return np.repeat(b, 3).astype(np.uint8)


def element_node_ids(mx, my, mz) -> np.ndarray:
    """[E, 27] node indices, element e = ei + mx*(ej + my*ek), local n = i+3j+9k."""
    X, Y = 2 * mx + 1, 2 * my + 1
    ek, ej, ei = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    k, j, i = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    return (2 * ei[:, None] + i[None]) + X * ((2 * ej[:, None] + j[None]) + Y * (2 * ek[:, None] + k[None]))


def _q2_basis_at_qps() -> np.ndarray:
    """[27 qp, 27 node] Q2 basis values at the 3x3x3 Gauss points."""
    x = GAUSS
    N1 = np.stack([0.5 * x * (x - 1.0), (1.0 + x) * (1.0 - x), 0.5 * (1.0 + x) * x], axis=1)  # [q, n]
    out = np.einsum("ai,bj,ck->cbakji", N1, N1, N1).reshape(27, 27)
    return out


def qp_coordinates(coords, mx, my, mz) -> np.ndarray:
    """[E, 27, 3] physical positions of the quadrature points."""
    nid = element_node_ids(mx, my, mz)
    xe = coords.reshape(-1, 3)[nid]  # [E, 27, 3]
    return np.einsum("qn,end->eqd", _q2_basis_at_qps(), xe)


def eta_loguniform(nel, lo=1e-2, hi=1e2, rng=None) -> np.ndarray:
    rng = rng or np.random.default_rng(SEED + 1)
    return 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), size=27 * nel)


def eta_sinker(qpx: np.ndarray, nspheres=32, rmin=0.05, rmax=0.12, eta_bg=1.0, eta_s=1e4,
               rng=None) -> np.ndarray:
    """Background eta_bg with eta_s inside seeded spheres, evaluated at qps."""
    rng = rng or np.random.default_rng(SEED + 2)
    cen = rng.uniform(0.0, 1.0, size=(nspheres, 3))
    rad = rng.uniform(rmin, rmax, size=nspheres)
    p = qpx.reshape(-1, 3)
    eta = np.full(p.shape[0], eta_bg)
    for c, r in zip(cen, rad):
        inside = ((p - c) ** 2).sum(axis=1) <= r * r
        eta[inside] = eta_s
    return eta


def eta_layered_ellipsoids(qpx: np.ndarray, n_ell=64, rng=None) -> np.ndarray:
    """Three z-layers at 1e3 / 1e1 / 1e-3 (bottom to top) plus seeded weak
    ellipsoids at 1e-3: viscosity contrast 1e6."""
    rng = rng or np.random.default_rng(SEED + 3)
    p = qpx.reshape(-1, 3)
    z = p[:, 2]
    eta = np.where(z < 1.0 / 3.0, 1e3, np.where(z < 2.0 / 3.0, 1e1, 1e-3))
    cen = rng.uniform(0.0, 1.0, size=(n_ell, 3))
    axes = rng.uniform(0.03, 0.1, size=(n_ell, 3))
    for c, a in zip(cen, axes):
        inside = (((p - c) / a) ** 2).sum(axis=1) <= 1.0
        eta[inside] = 1e-3
    return eta


def smooth_body_force(coords, mx, my, mz, rng=None) -> np.ndarray:
    """Smooth body-force load: f = a * (sin(pi x) sin(pi y), sin(pi y) sin(pi z),
    sin(pi z) sin(pi x)) with seeded amplitudes a_d in [0.5, 1.5], lumped to
    the nodes with the nodal volume 1 / (8 mx my mz); Dirichlet rows (all
    faces) zero."""
    rng = rng or np.random.default_rng(SEED + 5)
    a = rng.uniform(0.5, 1.5, size=3)
    c = coords.reshape(-1, 3)
    f = np.empty_like(c)
    sx, sy, sz = (np.sin(np.pi * c[:, d]) for d in range(3))
    np.multiply(sx, sy, out=f[:, 0])
    np.multiply(sy, sz, out=f[:, 1])
    np.multiply(sz, sx, out=f[:, 2])
    f *= a[None, :] / (8.0 * mx * my * mz)
    f = f.reshape(-1)
    f[dirichlet_all_faces(mx, my, mz).astype(bool)] = 0.0
    return f


def box_qp_axes(mx, my, mz):
    """Quadrature-point positions of the affine unit-cube box, per axis:
    gx[ei, a] = (ei + (1 + g_a) / 2) / mx, likewise gy, gz."""
    g = 0.5 * (1.0 + GAUSS)
    return tuple((np.arange(m)[:, None] + g[None, :]) / m for m in (mx, my, mz))


def _ranges(ax, c, r):
    """Element index range whose quadrature points may lie within [c-r, c+r]."""
    m = ax.shape[0]
    lo = int(np.clip(np.floor((c - r) * m) - 1, 0, m))
    hi = int(np.clip(np.ceil((c + r) * m) + 1, 0, m))
    return lo, hi


def eta_layered_ellipsoids_box(mx, my, mz, n_ell=64, rng=None) -> np.ndarray:
    """eta_layered_ellipsoids on the affine unit-cube box, evaluated per
    ellipsoid over its bounding box only (same seeded draws)."""
    rng = rng or np.random.default_rng(SEED + 3)
    gx, gy, gz = box_qp_axes(mx, my, mz)
    # element-major [27e + q] = array (ek, ej, ei, c, b, a)
    zl = np.where(gz < 1.0 / 3.0, 1e3, np.where(gz < 2.0 / 3.0, 1e1, 1e-3))  # (mz, 3)
    eta = np.empty((mz, my, mx, 3, 3, 3))
    eta[...] = zl[:, None, None, :, None, None]
    cen = rng.uniform(0.0, 1.0, size=(n_ell, 3))
    axes = rng.uniform(0.03, 0.1, size=(n_ell, 3))
    for c, a in zip(cen, axes):
        i0, i1 = _ranges(gx, c[0], a[0])
        j0, j1 = _ranges(gy, c[1], a[1])
        k0, k1 = _ranges(gz, c[2], a[2])
        dx = ((gx[i0:i1] - c[0]) / a[0]) ** 2  # (ni, 3a)
        dy = ((gy[j0:j1] - c[1]) / a[1]) ** 2
        dz = ((gz[k0:k1] - c[2]) / a[2]) ** 2
        r2 = (dz[:, None, None, :, None, None] + dy[None, :, None, None, :, None]
              + dx[None, None, :, None, None, :])
        sub = eta[k0:k1, j0:j1, i0:i1]
        sub[r2 <= 1.0] = 1e-3
    return eta.reshape(-1)


def random_velocity(mx, my, mz, bcmask=None, rng=None) -> np.ndarray:
    rng = rng or np.random.default_rng(SEED + 4)
    u = rng.uniform(-1.0, 1.0, size=3 * nnodes(mx, my, mz))
    if bcmask is not None:
        u[bcmask.astype(bool)] = 0.0
    return u


def reference_test_vector(coords) -> np.ndarray:
    """f = x + y + z + component + 3 (test_stokes_operators.c:48-92)."""
    c = coords.reshape(-1, 3)
    s = c.sum(axis=1)
    return np.stack([s + 3.0, s + 4.0, s + 5.0], axis=1).reshape(-1)


# ---- named workloads -------------------------------------------------------

def config1(m=16, seed=SEED):
    """16^3 perturbed unit cube, log-uniform eta in [1e-2,1e2], random u."""
    rng = np.random.default_rng(seed)
    coords = perturb_interior(box_coords(m, m, m), m, m, m, 0.1, rng)
    eta = eta_loguniform(m ** 3, 1e-2, 1e2, rng)
    bc = dirichlet_all_faces(m, m, m)
    u = random_velocity(m, m, m, bc, rng)
    return dict(mx=m, my=m, mz=m, coords=coords, eta=eta, bc=bc, u=u)


def config2(m=64, seed=SEED):
    """64^3 affine unit cube, sinker viscosity (32 spheres at 1e4), random u."""
    rng = np.random.default_rng(seed)
    coords = box_coords(m, m, m)
    eta = eta_sinker(qp_coordinates(coords, m, m, m), rng=rng)
    bc = dirichlet_all_faces(m, m, m)
    u = random_velocity(m, m, m, bc, rng)
    return dict(mx=m, my=m, mz=m, coords=coords, eta=eta, bc=bc, u=u)


def config3(m=128, seed=SEED):
    """128^3 unit cube, layered viscosity 1e3/1e1/1e-3 with 64 weak ellipsoids
    (contrast 1e6), smooth body-force load."""
    rng = np.random.default_rng(seed)
    coords = box_coords(m, m, m)
    eta = eta_layered_ellipsoids_box(m, m, m, rng=rng)
    bc = dirichlet_all_faces(m, m, m)
    b = smooth_body_force(coords, m, m, m, rng)
    return dict(mx=m, my=m, mz=m, coords=coords, eta=eta, bc=bc, b=b)


# ---- saddle-point rift workload ---------------------------------------------

RIFT_LEN = (2.0, 2.0, 1.0)  # domain [0,2] x [0,2] x [0,1], z up


def rift_bc(mx, my, mz):
    """Free-slip sides and base, extension on the x-faces, free top surface:
    u_x fixed on x = 0 and x = 2 (values -1 and +1), u_y = 0 on the y-faces,
    u_z = 0 on the base.  Returns the 3N Dirichlet mask and the 3N vector of
    boundary values (zero off the boundary)."""
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    m = np.zeros((Z, Y, X, 3), dtype=np.uint8)
    g = np.zeros((Z, Y, X, 3))
    m[:, :, 0, 0] = m[:, :, -1, 0] = 1
    g[:, :, 0, 0], g[:, :, -1, 0] = -1.0, 1.0
    m[:, 0, :, 1] = m[:, -1, :, 1] = 1
    m[0, :, :, 2] = 1
    return m.reshape(-1), g.reshape(-1)


def rift_prefactor(mx, my, mz, nexp=3.0, rng=None) -> np.ndarray:
    """Power-law prefactor B at the quadrature points (element-major 27E):
    a crust (z > 0.75, eta 1e3) over a mantle (eta 1), and a seeded weak notch
    (eta 1e-3) at the base of the crust across the whole y extent; B = eta *
    e_ref^(1 - 1/n) with the reference strain rate e_ref = 1 of the imposed
    extension, so eta(u) equals the layer value at e_ref."""
    rng = rng or np.random.default_rng(SEED + 7)
    gx, gy, gz = box_qp_axes(mx, my, mz)
    gx, gy, gz = gx * RIFT_LEN[0], gy * RIFT_LEN[1], gz * RIFT_LEN[2]
    xc = rng.uniform(0.9, 1.1)
    hw = rng.uniform(0.04, 0.08)
    zl = np.where(gz > 0.75, 1e3, 1.0)  # (mz, 3)
    notch_x = np.abs(gx - xc) <= hw  # (mx, 3)
    notch_z = (gz > 0.75) & (gz < 0.85)  # (mz, 3)
    eta = np.empty((mz, my, mx, 3, 3, 3))
    eta[...] = zl[:, None, None, :, None, None]
    inside = notch_z[:, None, None, :, None, None] & notch_x[None, None, :, None, None, :]
    eta[np.broadcast_to(inside, eta.shape)] = 1e-3
    return eta.reshape(-1)  # e_ref = 1: B = eta


def rift_initial_velocity(coords, rng=None) -> np.ndarray:
    """Seeded initial velocity for the strain-rate-dependent viscosity: the
    divergence-free extension u = (x - 1, 0, -z) plus a seeded smooth
    perturbation of amplitude 0.2."""
    rng = rng or np.random.default_rng(SEED + 8)
    c = coords.reshape(-1, 3)
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    k = rng.integers(1, 4, size=(3, 3))
    ph = rng.uniform(0, 2 * np.pi, size=(3, 3))
    u = np.stack([x - 1.0, np.zeros_like(x), -z], axis=1)
    for d in range(3):
        u[:, d] += 0.2 * (np.sin(np.pi * k[d, 0] * x / 2 + ph[d, 0]) * np.sin(np.pi * k[d, 1] * y / 2 + ph[d, 1])
                          * np.sin(np.pi * k[d, 2] * z + ph[d, 2]))
    return u.reshape(-1)


def rift_body_force(coords, mx, my, mz) -> np.ndarray:
    """Gravity load -rho g e_z (rho 2.7 in the crust z > 0.75, 3.3 below,
    g = 1) lumped to the nodes with the nodal volume of the box."""
    c = coords.reshape(-1, 3)
    f = np.zeros_like(c)
    vol = RIFT_LEN[0] * RIFT_LEN[1] * RIFT_LEN[2] / (8.0 * mx * my * mz)
    f[:, 2] = -np.where(c[:, 2] > 0.75, 2.7, 3.3) * vol
    return f.reshape(-1)


def config5(mx=192, my=192, mz=96, seed=SEED):
    """Rift saddle point: 192 x 192 x 96 elements on [0,2]x[0,2]x[0,1]."""
    rng = np.random.default_rng(seed)
    coords = box_coords(mx, my, mz, hi=RIFT_LEN)
    B = rift_prefactor(mx, my, mz, rng=rng)
    u0 = rift_initial_velocity(coords, rng)
    bc, g = rift_bc(mx, my, mz)
    f = rift_body_force(coords, mx, my, mz)
    return dict(mx=mx, my=my, mz=mz, coords=coords, B=B, u0=u0, bc=bc, g=g, f=f, nexp=3.0,
                eta_min=1e-3, eta_max=1e3)


# ---- the reference's viscous_sinker operator test (8 x 9 x 11) ---------------

def gaussian_bump_coords(mx, my, mz, amp=-0.3, lam=-5.6) -> np.ndarray:
    """Unit box whose top surface (y = 1) is lowered by a Gaussian bump:
    on [-1,1]^3, y_j = -1 + j (h(x,z) + 1) / (Y - 1) with h = amp exp(lam (x^2 +
    z^2)) + 1, then mapped back to [0,1]^3 (MeshDeformation_GaussianBump_YMAX,
    mesh_deformation.c:48-130, as the viscous_sinker model applies it,
    model_viscous_sinker_ops.c:668-680)."""
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    xs = -1.0 + np.arange(X) * (2.0 / (X - 1))
    zs = -1.0 + np.arange(Z) * (2.0 / (Z - 1))
    xn, zn = np.meshgrid(xs, zs)  # (Z, X)
    dy = (amp * np.exp(lam * (xn * xn + zn * zn)) + 1.0 + 1.0) / (Y - 1)
    yn = -1.0 + np.arange(Y)[None, :, None] * dy[:, None, :]  # (Z, Y, X)
    c = np.empty((Z, Y, X, 3))
    c[..., 0] = ((xn + 1.0) / 2.0)[:, None, :]
    c[..., 1] = (yn + 1.0) / 2.0
    c[..., 2] = ((zn + 1.0) / 2.0)[:, None, :]
    return c.reshape(-1)


def _q2_1d(t):
    return np.stack([0.5 * t * (t - 1.0), (1.0 + t) * (1.0 - t), 0.5 * (1.0 + t) * t], axis=-1)


def sinker_markers(coords, mx, my, mz, eta0=1.0, eta1=10.0, nxp=2):
    """Material points of the viscous_sinker model: an unperturbed nxp^3
    lattice per element at xi = -1 + (2/nxp)(i + 1/2)
    (SwarmMPntStd_CoordAssignment_LatticeLayout3d, material_point_std_utils.c:
    100-210), viscosity eta1 inside the sphere of radius 0.25 about the box
    centre and eta0 outside (single spherical inclusion,
    model_viscous_sinker_ops.c:715-790).  Returns (element, xi [n,3], eta)."""
    g = -1.0 + (2.0 / nxp) * (np.arange(nxp) + 0.5)
    xi = np.array([[g[a], g[b], g[c]] for c in range(nxp) for b in range(nxp) for a in range(nxp)])
    npc = xi.shape[0]
    N = np.einsum("pi,pj,pk->pkji", _q2_1d(xi[:, 0]), _q2_1d(xi[:, 1]), _q2_1d(xi[:, 2])).reshape(npc, 27)
    nid = element_node_ids(mx, my, mz)
    xe = coords.reshape(-1, 3)[nid]
    xp = np.einsum("pn,end->epd", N, xe)
    inside = (((xp - 0.5) / 0.25) ** 2).sum(axis=-1) < 1.0
    E = mx * my * mz
    elem = np.repeat(np.arange(E), npc)
    return elem, np.tile(xi, (E, 1)), np.where(inside, eta1, eta0).reshape(-1)


def sinker_bc(mx, my, mz) -> np.ndarray:
    """Free slip with a free top (VSBC_FreeSlipFreeSurface, the viscous_sinker
    default, model_viscous_sinker_ops.c:236-244): u_x on the x-faces, u_y on
    the base y = 0, u_z on the z-faces; 3N mask."""
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    m = np.zeros((Z, Y, X, 3), dtype=np.uint8)
    m[:, :, 0, 0] = m[:, :, -1, 0] = 1
    m[:, 0, :, 1] = 1
    m[0, :, :, 2] = m[-1, :, :, 2] = 1
    return m.reshape(-1)


def reference_pressure_test_vector(mx, my, mz) -> np.ndarray:
    """p[4e+i] = X + Y + Z + i + 3 with X = -1 + ei 2/(mx-1) etc.
    (_GenerateTestVectorDAP, test_stokes_operators.c:94-136)."""
    ei, ej, ek = np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij")
    e = (ei + mx * (ej + my * ek)).ravel()
    base = (-1.0 + ei * 2.0 / (mx - 1) - 1.0 + ej * 2.0 / (my - 1) - 1.0 + ek * 2.0 / (mz - 1)).ravel()
    p = np.zeros(4 * mx * my * mz)
    for i in range(4):
        p[4 * e + i] = base + i + 3.0
    return p
