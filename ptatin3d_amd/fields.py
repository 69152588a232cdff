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
