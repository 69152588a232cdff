"""Host-side mirror of the reference's GMG-preconditioned viscous solve.

`GMG` holds a level hierarchy of matrix-free A11 operators (level 0 =
coarsest, as PETSc PCMG) on the device, configured like the reference's
options (pTatin3dStokesKSPConfigureFSGMG, ptatin_driver_linear_ts.c:819-882):
Chebyshev/Jacobi smoothing with eigenvalue bounds from an estimate
(-..._esteig a,b,c,d), transfers from DMDA Q1 lattice interpolation, an exact
coarse solve and FGMRES outside.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, fields as F
from ._lib import DeviceArray, MGOptions, call, require


def default_options(**kw) -> MGOptions:
    o = MGOptions()
    call("ptmf_mg_default_options", C.byref(o))
    for k, v in kw.items():
        if k == "esteig":
            for i in range(4):
                o.esteig[i] = float(v[i])
        else:
            setattr(o, k, v)
    return o


class GMG:
    def __init__(self, mx: int, my: int, mz: int, nlevels: int):
        h = C.c_void_p()
        call("ptmf_mg_create", int(mx), int(my), int(mz), int(nlevels), C.byref(h))
        self.handle = h
        self.nlevels = nlevels
        self.dims = [(mx >> (nlevels - 1 - l), my >> (nlevels - 1 - l), mz >> (nlevels - 1 - l))
                     for l in range(nlevels)]
        self.n = 3 * F.nnodes(mx, my, mz)
        self.nstokes = self.n + 4 * mx * my * mz

    def level_n(self, level: int) -> int:
        """velocity dofs of a level (3 * nodes)"""
        if not 0 <= int(level) < self.nlevels:
            raise _lib.PtmfError(1, "GMG", f"level {level} outside 0..{self.nlevels - 1}")
        return 3 * F.nnodes(*self.dims[int(level)])

    def level_coordinates(self, fine_coords: np.ndarray) -> list[np.ndarray]:
        """Injected coordinates of every level (for host-side field evaluation)."""
        out = [None] * self.nlevels
        cur = np.asarray(fine_coords, dtype=np.float64).reshape(-1, 3)
        for l in range(self.nlevels - 1, -1, -1):
            mx, my, mz = self.dims[l]
            out[l] = cur.reshape(-1)
            if l == 0:
                break
            X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
            c3 = cur.reshape(Z, Y, X, 3)
            cur = np.ascontiguousarray(c3[::2, ::2, ::2]).reshape(-1, 3)
        return out

    def set_coordinates(self, coords: np.ndarray) -> None:
        c = np.ascontiguousarray(coords, dtype=np.float64)
        call("ptmf_mg_set_coordinates", self.handle, _lib._hp(c))

    def set_viscosity(self, level: int, eta: np.ndarray) -> None:
        e = np.ascontiguousarray(eta, dtype=np.float64)
        call("ptmf_mg_set_viscosity", self.handle, int(level), _lib._hp(e))

    def coarsen_viscosity(self, mode: str = "arithmetic") -> None:
        modes = {"arithmetic": 0, "geometric": 1, "harmonic": 2, "qp-arithmetic": 3, "qp-geometric": 4,
                 "q1-projection": 5, "q1-injection": 6}
        call("ptmf_mg_coarsen_viscosity", self.handle, modes[mode])

    def set_dirichlet(self, level: int, mask: np.ndarray) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        call("ptmf_mg_set_dirichlet", self.handle, int(level), _lib._hp(m))

    def set_dirichlet_all_faces(self) -> None:
        for l, (mx, my, mz) in enumerate(self.dims):
            self.set_dirichlet(l, F.dirichlet_all_faces(mx, my, mz))

    def setup(self, opt: MGOptions | None = None) -> None:
        call("ptmf_mg_setup", self.handle, C.byref(opt) if opt is not None else None)
        self.opt = opt if opt is not None else default_options()

    def coarse_info(self) -> dict:
        """The coarse factor of the last set-up (dense inverse or sparse
        multifrontal Cholesky), its bytes, tree depth and factorisation ms."""
        k, b, d, ms = C.c_int(), C.c_long(), C.c_int(), C.c_double()
        call("ptmf_mg_coarse_info", self.handle, C.byref(k), C.byref(b), C.byref(d), C.byref(ms))
        return {"kind": {1: "dense", 2: "sparse"}[k.value], "bytes": b.value, "depths": d.value,
                "setup_ms": ms.value}

    def eigs(self, level: int) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        call("ptmf_mg_level_eigs", self.handle, int(level), C.byref(a), C.byref(b))
        return a.value, b.value

    def set_eigs(self, level: int, emin: float, emax: float) -> None:
        call("ptmf_mg_set_level_eigs", self.handle, int(level), float(emin), float(emax))

    def level_op_handle(self, level: int) -> C.c_void_p:
        h = C.c_void_p()
        call("ptmf_mg_level_op", self.handle, int(level), C.byref(h))
        return h

    def level_apply(self, level: int, x: DeviceArray, y: DeviceArray) -> None:
        prec = 32 if x.dtype == np.float32 else 64
        require("GMG.level_apply", self.level_n(level), x.dtype, x, y)
        call("ptmf_mg_level_apply", self.handle, int(level), prec, x.ptr, y.ptr)

    def level_assemble(self, level: int):
        """The level's operator as host CSR arrays (rowptr, colidx, vals) in
        the node-interlaced dof order, the reference's assembled operator
        (MatAssemble_StokesA_AUU, stokes_assembly.c:115-297, Dirichlet rows
        and columns replaced by a unit diagonal, dmda_bcs.c:1035-1060)."""
        n = self.level_n(level)
        nnz = C.c_long()
        rowptr = np.zeros(n + 1, dtype=np.int64)
        call("ptmf_mg_level_assemble", self.handle, int(level), C.byref(nnz), _lib._hp(rowptr), None, None)
        colidx = np.zeros(nnz.value, dtype=np.int64)
        vals = np.zeros(nnz.value, dtype=np.float64)
        call("ptmf_mg_level_assemble", self.handle, int(level), C.byref(nnz), _lib._hp(rowptr), _lib._hp(colidx), _lib._hp(vals))
        return rowptr, colidx, vals

    def level_rigid_body_modes(self, level: int) -> np.ndarray:
        """The six rigid-body modes of the level from its coordinates, shape
        (6, 3N): translations x, y, z, rotations about z, x, y (the near null
        space of ptatin_driver_linear_ts.c:663-666)."""
        n = self.level_n(level)
        modes = np.zeros((6, n), dtype=np.float64)
        call("ptmf_mg_level_rigid_body_modes", self.handle, int(level), _lib._hp(modes))
        return modes

    def vcycle(self, b: DeviceArray, x: DeviceArray) -> None:
        require("GMG.vcycle", self.n, np.float64, b, x)
        call("ptmf_mg_vcycle", self.handle, b.ptr, x.ptr)

    def prolong_add(self, level: int, xc: DeviceArray, yf: DeviceArray) -> None:
        require("GMG.prolong_add", self.level_n(level - 1), np.float64, xc)
        require("GMG.prolong_add", self.level_n(level), np.float64, yf)
        call("ptmf_mg_prolong_add", self.handle, int(level), xc.ptr, yf.ptr)

    def restrict(self, level: int, xf: DeviceArray, yc: DeviceArray) -> None:
        require("GMG.restrict", self.level_n(level), np.float64, xf)
        require("GMG.restrict", self.level_n(level - 1), np.float64, yc)
        call("ptmf_mg_restrict", self.handle, int(level), xf.ptr, yc.ptr)

    def smooth(self, level: int, b: DeviceArray, x: DeviceArray, its: int, nonzero_guess: bool) -> None:
        prec = 32 if b.dtype == np.float32 else 64
        require("GMG.smooth", self.level_n(level), b.dtype, b, x)
        call("ptmf_mg_smooth", self.handle, int(level), prec, int(its), int(nonzero_guess), b.ptr, x.ptr)

    def fgmres(self, b: DeviceArray, x: DeviceArray, rtol=1e-8, maxit=200, restart=30,
               history_len=512):
        require("GMG.fgmres", self.n, np.float64, b, x)
        its = C.c_int(0)
        rn = C.c_double(0)
        hist = np.zeros(history_len)
        call("ptmf_mg_fgmres", self.handle, b.ptr, x.ptr, float(rtol), int(maxit), int(restart),
             C.byref(its), C.byref(rn), hist.ctypes.data_as(C.POINTER(C.c_double)), int(history_len))
        nz = np.flatnonzero(hist)
        return its.value, rn.value, hist[: (nz[-1] + 1) if nz.size else 0]

    def fcg(self, b: DeviceArray, x: DeviceArray, rtol=1e-8, maxit=500, history_len=1024):
        """Flexible CG preconditioned by the V-cycle, from the given x."""
        require("GMG.fcg", self.n, np.float64, b, x)
        its = C.c_int(0)
        rn = C.c_double(0)
        hist = np.zeros(history_len)
        call("ptmf_mg_fcg", self.handle, b.ptr, x.ptr, float(rtol), int(maxit),
             C.byref(its), C.byref(rn), hist.ctypes.data_as(C.POINTER(C.c_double)), int(history_len))
        nz = np.flatnonzero(hist)
        return its.value, rn.value, hist[: (nz[-1] + 1) if nz.size else 0]

    # -- saddle point -------------------------------------------------------
    def stokes_setup(self, inner_its: int = 1, inner_rtol: float = 0.0) -> None:
        """Saddle-point preconditioner; A11^-1 is one V-cycle (inner_its <= 1)
        or up to inner_its V-cycle-preconditioned flexible-CG iterations."""
        call("ptmf_mg_stokes_setup", self.handle, int(inner_its), float(inner_rtol))

    def set_schur(self, kind: str) -> None:
        """Schur approximation of the saddle-point preconditioner: "mass"
        (inverse-viscosity-scaled pressure mass, the default) or "bdb"
        (element blocks of B^T diag(A11)^-1 B).  Call after stokes_setup."""
        call("ptmf_mg_stokes_set_schur", self.handle, {"mass": 0, "bdb": 1}[kind])

    def stokes_apply(self, x: DeviceArray, y: DeviceArray) -> None:
        require("GMG.stokes_apply", self.nstokes, np.float64, x, y)
        call("ptmf_mg_stokes_apply", self.handle, x.ptr, y.ptr)

    def stokes_fgmres(self, b: DeviceArray, x: DeviceArray, rtol=1e-8, maxit=500, restart=50,
                      history_len=1024):
        require("GMG.stokes_fgmres", self.nstokes, np.float64, b, x)
        its = C.c_int(0)
        rn = C.c_double(0)
        hist = np.zeros(history_len)
        call("ptmf_mg_stokes_fgmres", self.handle, b.ptr, x.ptr, float(rtol), int(maxit), int(restart),
             C.byref(its), C.byref(rn), hist.ctypes.data_as(C.POINTER(C.c_double)), int(history_len))
        nz = np.flatnonzero(hist)
        return its.value, rn.value, hist[: (nz[-1] + 1) if nz.size else 0]

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            call("ptmf_mg_destroy", C.byref(self.handle))
        self.handle = None

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value and _lib._lib is not None:
                _lib._lib.ptmf_mg_destroy(C.byref(self.handle))
        except Exception:
            pass


def solve_timed(mg: GMG, b: np.ndarray, rtol=1e-8, maxit=200, restart=30, krylov="fgmres", x0=None):
    """Upload b, run FGMRES (or flexible CG) from x0 (default 0); returns (x,
    its, rnorm, history, seconds)."""
    db = DeviceArray.from_host(b)
    if x0 is None:
        dx = DeviceArray(b.size)
        dx.zero()
    else:
        dx = DeviceArray.from_host(np.asarray(x0, dtype=np.float64))
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    if krylov == "fcg":
        its, rn, hist = mg.fcg(db, dx, rtol, maxit)
    else:
        its, rn, hist = mg.fgmres(db, dx, rtol, maxit, restart)
    call("ptmf_device_synchronize")
    t1 = time.perf_counter()
    x = dx.download()
    db.free()
    dx.free()
    return x, its, rn, hist, t1 - t0


def stokes_solve_timed(mg: GMG, bu: np.ndarray, bp: np.ndarray, rtol=1e-8, maxit=500, restart=50,
                       history_len=1024, x0=None):
    """Saddle-point solve from x0 = [u | p] (default 0); returns (u, p, its,
    rnorm, history, seconds)."""
    nu = bu.size
    b = np.concatenate([bu, bp])
    db = DeviceArray.from_host(b)
    if x0 is None:
        dx = DeviceArray(b.size)
        dx.zero()
    else:
        dx = DeviceArray.from_host(np.asarray(x0, dtype=np.float64))
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    its, rn, hist = mg.stokes_fgmres(db, dx, rtol, maxit, restart, history_len)
    call("ptmf_device_synchronize")
    t1 = time.perf_counter()
    x = dx.download()
    db.free()
    dx.free()
    return x[:nu], x[nu:], its, rn, hist, t1 - t0


# -- workloads with non-homogeneous boundary values ----------------------------

def powerlaw_viscosity(dims, coords, u, B, nexp, eta_min, eta_max) -> np.ndarray:
    """Quadrature-point power-law viscosity of velocity u, evaluated on the
    device (element-major 27E, host copy)."""
    from .operators import MatA11MF
    mx, my, mz = dims
    E = mx * my * mz
    A = MatA11MF(mx, my, mz).setup(coords, np.ones(27 * E), np.zeros(3 * F.nnodes(mx, my, mz), np.uint8))
    du, dB, de = DeviceArray.from_host(u), DeviceArray.from_host(B), DeviceArray(27 * E)
    A.powerlaw_viscosity(du, dB, nexp, eta_min, eta_max, de)
    eta = de.download()
    for d in (du, dB, de):
        d.free()
    A.destroy()
    return eta


def stokes_rhs_lifted(dims, coords, eta, bc, g, f):
    """Right-hand side of K [u|p] = b for Dirichlet values g on the masked
    dofs: b_u = f - A11 g, b_p = -A21 g with the unmasked operators, then
    b_u = g on the Dirichlet dofs (the rows of K there are the identity)."""
    from .operators import MatStokesMF
    mx, my, mz = dims
    A = MatStokesMF(mx, my, mz).setup(coords, eta, np.zeros_like(bc))
    yu, yp = A.stokes_host(g, np.zeros(4 * mx * my * mz))
    A.destroy()
    m = np.asarray(bc).astype(bool)
    bu = f - yu
    bu[m] = g[m]
    return bu, -yp


def rift_hierarchy(c, nlevels, eta, opt: MGOptions, coarsen="qp-arithmetic", inner_its=1, inner_rtol=0.0,
                   schur="mass"):
    """GMG hierarchy of the rift saddle point (free-slip sides and base,
    x-face extension on every level) with the saddle-point preconditioner."""
    mg = GMG(c["mx"], c["my"], c["mz"], nlevels)
    mg.set_coordinates(c["coords"])
    mg.set_viscosity(nlevels - 1, eta)
    mg.coarsen_viscosity(coarsen)
    for l, d in enumerate(mg.dims):
        mg.set_dirichlet(l, F.rift_bc(*d)[0])
    mg.setup(opt)
    mg.stokes_setup(inner_its, inner_rtol)
    if schur != "mass":
        mg.set_schur(schur)
    return mg


class BatchGMG:
    """Velocity-block solves of many independent boxes in lockstep: one launch
    per stage over all boxes, the V-cycle captured once as a HIP graph.  The
    GMG objects (set up, same levels and smoother options) must outlive it."""

    def __init__(self, mgs):
        self.mgs = list(mgs)
        arr = (C.c_void_p * len(self.mgs))(*[g.handle.value for g in self.mgs])
        h = C.c_void_p()
        call("ptmf_bmg_create", arr, len(self.mgs), C.byref(h))
        self.handle = h

    @staticmethod
    def _ptrs(arrs):
        return (C.c_void_p * len(arrs))(*[a.ptr for a in arrs])

    def _check(self, bs, xs):
        if len(bs) != len(self.mgs) or len(xs) != len(self.mgs):
            raise _lib.PtmfError(1, "BatchGMG", f"{len(bs)}/{len(xs)} arrays for {len(self.mgs)} boxes")
        for g, b, x in zip(self.mgs, bs, xs):
            require("BatchGMG", g.n, np.float64, b, x)

    def vcycle(self, bs, xs) -> None:
        self._check(bs, xs)
        call("ptmf_bmg_vcycle", self.handle, self._ptrs(bs), self._ptrs(xs))

    def fcg(self, bs, xs, rtol=1e-8, maxit=500):
        self._check(bs, xs)
        n = len(self.mgs)
        its = (C.c_int * n)()
        rn = (C.c_double * n)()
        call("ptmf_bmg_fcg", self.handle, self._ptrs(bs), self._ptrs(xs), float(rtol), int(maxit), its, rn)
        return np.array(its[:]), np.array(rn[:])

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            call("ptmf_bmg_destroy", C.byref(self.handle))
        self.handle = None

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value and _lib._lib is not None:
                _lib._lib.ptmf_bmg_destroy(C.byref(self.handle))
        except Exception:
            pass
