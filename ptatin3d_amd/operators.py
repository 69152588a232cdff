"""Host-side mirror of pTatin3d's matrix-free operator objects over the ptmf C-ABI.

`MatA11MF` plays the role of the reference's `struct _p_MatA11MF`
(include/stokes_operators.h:47-62) with the `-a11_op hip` backend:
create -> setup (coordinates, quadrature viscosity, Dirichlet list) ->
MatMult / MatGetDiagonal.  `MatStokesMF` adds the B / B^T couplings
(MatMult_MFStokes_A12 / _A21).  Arrays are numpy on the host or
`DeviceArray` on the GPU; layouts are those of include/ptmf.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, require

A11_OPERATOR_TYPES = ("hip",)


class MatA11MF:
    """Matrix-free Q2 viscous block on one structured box, device resident."""

    def __init__(self, mx: int, my: int, mz: int, a11_op: str = "hip"):
        # MatA11MFCreate (stokes_operators.c:101-157): unknown backends are an error
        if a11_op not in A11_OPERATOR_TYPES:
            raise _lib.PtmfError(5, "MatA11MFCreate", f"No -a11_op {a11_op}")
        self.mx, self.my, self.mz = int(mx), int(my), int(mz)
        h = C.c_void_p()
        call("ptmf_op_create", self.mx, self.my, self.mz, C.byref(h))
        self.handle = h
        self.nnodes = (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)
        self.nel = mx * my * mz
        self.is_setup = False

    @classmethod
    def borrow(cls, handle, mx: int, my: int, mz: int) -> "MatA11MF":
        """A view of an operator owned elsewhere (a multigrid level's, from
        ptmf_mg_level_op); destroy() and garbage collection leave it alone."""
        op = cls.__new__(cls)
        op.mx, op.my, op.mz = int(mx), int(my), int(mz)
        op.handle = handle
        op.nnodes = (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)
        op.nel = mx * my * mz
        op.is_setup = True
        op._owned = False
        return op

    # -- MatA11MFSetup ----------------------------------------------------
    def setup(self, coords: np.ndarray, eta: np.ndarray, bcmask: np.ndarray) -> "MatA11MF":
        self.set_coordinates(coords)
        self.set_viscosity(eta)
        self.set_dirichlet(bcmask)
        self.is_setup = True
        return self

    def set_coordinates(self, coords: np.ndarray) -> None:
        c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        assert c.size == 3 * self.nnodes, (c.size, self.nnodes)
        call("ptmf_op_set_coordinates", self.handle, _lib._hp(c))

    def set_viscosity(self, eta: np.ndarray) -> None:
        e = np.ascontiguousarray(eta, dtype=np.float64).reshape(-1)
        assert e.size == 27 * self.nel
        call("ptmf_op_set_viscosity", self.handle, _lib._hp(e))

    def set_dirichlet(self, bcmask: np.ndarray) -> None:
        m = np.ascontiguousarray(bcmask, dtype=np.uint8).reshape(-1)
        assert m.size == 3 * self.nnodes
        call("ptmf_op_set_dirichlet", self.handle, _lib._hp(m))

    def set_tiling(self, TX: int, TY: int, ZC: int) -> None:
        """Element tile (TX*TY <= 21) and z-chunk of the MatMult kernel."""
        call("ptmf_op_set_tiling", self.handle, int(TX), int(TY), int(ZC))

    def get_tiling(self) -> tuple[int, int, int]:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("ptmf_op_get_tiling", self.handle, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def set_kernel_variant(self, variant: int) -> None:
        """0: z-march; 1: z-march with next-layer loads issued one stage ahead;
        2: stored quadrature-point geometry; 3: coloured scatter (no tiles)."""
        call("ptmf_op_set_kernel_variant", self.handle, int(variant))

    def set_tiles_per_group(self, tiles: int) -> None:
        """Tiles (one wave each) per work-group of the variant-0 MatMult: 1, 2 or 4."""
        call("ptmf_op_set_tiles_per_group", self.handle, int(tiles))

    def get_tiles_per_group(self) -> int:
        v = C.c_int()
        call("ptmf_op_get_tiles_per_group", self.handle, C.byref(v))
        return v.value

    def get_kernel_variant(self) -> int:
        v = C.c_int()
        call("ptmf_op_get_kernel_variant", self.handle, C.byref(v))
        return v.value

    @property
    def state(self) -> int:
        s = C.c_ulong(0)
        call("ptmf_op_state", self.handle, C.byref(s))
        return s.value

    # -- operations on device arrays --------------------------------------
    def mult(self, x: DeviceArray, y: DeviceArray) -> None:
        """MatMult_MFStokes_A11 (stokes_operators.c:1142-1205)."""
        require("MatA11MF.mult", 3 * self.nnodes, x.dtype, x, y)
        if x.dtype == np.float32:
            call("ptmf_a11_mult_f32", self.handle, x.ptr, y.ptr)
        else:
            call("ptmf_a11_mult", self.handle, x.ptr, y.ptr)
        # the C entry point is asynchronous; other operators run on their own
        # streams, so the Python host waits before y is used elsewhere
        call("ptmf_op_synchronize", self.handle)

    def spmv(self, u: DeviceArray, y: DeviceArray) -> None:
        """SpMVOp_MatMult: y += A11_local u (no boundary handling)."""
        require("MatA11MF.spmv", 3 * self.nnodes, np.float64, u, y)
        call("ptmf_a11_spmv", self.handle, u.ptr, y.ptr)

    def get_diagonal(self, d: DeviceArray) -> None:
        """MatGetDiagonal_MFStokes_A11 (stokes_operators.c:1060-1099)."""
        require("MatA11MF.get_diagonal", 3 * self.nnodes, np.float64, d)
        call("ptmf_a11_diagonal", self.handle, d.ptr)

    def powerlaw_viscosity(self, u: DeviceArray, B: DeviceArray, nexp: float, eta_min: float,
                           eta_max: float, eta: DeviceArray) -> None:
        """Quadrature-point power-law viscosity of velocity u (element-major 27E)."""
        require("MatA11MF.powerlaw_viscosity", 3 * self.nnodes, np.float64, u)
        require("MatA11MF.powerlaw_viscosity", 27 * self.nel, np.float64, B, eta)
        call("ptmf_op_powerlaw_viscosity", self.handle, u.ptr, B.ptr, float(nexp), float(eta_min),
             float(eta_max), eta.ptr)

    def project_points(self, elem: DeviceArray, xi: DeviceArray, val: DeviceArray, eta: DeviceArray) -> None:
        """Material-point values to the quadrature points (element-major 27E)."""
        require("MatA11MF.project_points", val.n, np.int64, elem)
        require("MatA11MF.project_points", 3 * val.n, np.float64, xi)
        require("MatA11MF.project_points", 27 * self.nel, np.float64, eta)
        call("ptmf_op_project_points", self.handle, int(val.n), elem.ptr, xi.ptr, val.ptr, eta.ptr)

    def set_viscosity_device(self, eta: DeviceArray) -> None:
        require("MatA11MF.set_viscosity_device", 27 * self.nel, np.float64, eta)
        call("ptmf_op_set_viscosity_device", self.handle, eta.ptr)

    def get_l1_diagonal(self, d: DeviceArray) -> None:
        """l1 row sums of the MatMult operator (the l1-Jacobi scaling)."""
        require("MatA11MF.get_l1_diagonal", 3 * self.nnodes, np.float64, d)
        call("ptmf_a11_l1_diagonal", self.handle, d.ptr)

    # -- host conveniences --------------------------------------------------
    def mult_host(self, x: np.ndarray, precision: int = 64) -> np.ndarray:
        dt = np.float64 if precision == 64 else np.float32
        dx = DeviceArray.from_host(np.asarray(x, dtype=dt))
        dy = DeviceArray(dx.n, dt)
        self.mult(dx, dy)
        out = dy.download()
        dx.free()
        dy.free()
        return out

    def spmv_host(self, u: np.ndarray, Yu: np.ndarray) -> None:
        """MFStokesWrapper_A11_HIP contract on host arrays: Yu += A u."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 3 * self.nnodes or Yu.size != 3 * self.nnodes:
            raise _lib.PtmfError(1, "MatA11MF.spmv_host", f"{u.size}/{Yu.size} values, expected {3 * self.nnodes}")
        if Yu.dtype != np.float64 or not Yu.flags.c_contiguous:
            raise _lib.PtmfError(1, "MatA11MF.spmv_host", "Yu must be a contiguous float64 array")
        call("ptmf_a11_spmv_host", self.handle, _lib._hp(u), _lib._hp(Yu))

    def l1_diagonal_host(self) -> np.ndarray:
        d = DeviceArray(3 * self.nnodes)
        self.get_l1_diagonal(d)
        out = d.download()
        d.free()
        return out

    def diagonal_host(self) -> np.ndarray:
        d = DeviceArray(3 * self.nnodes)
        self.get_diagonal(d)
        out = d.download()
        d.free()
        return out

    def time_mult(self, x: DeviceArray, y: DeviceArray, reps: int) -> float:
        """ms per MatMult over `reps` back-to-back applies (HIP events)."""
        require("MatA11MF.time_mult", 3 * self.nnodes, x.dtype, x, y)
        prec = 32 if x.dtype == np.float32 else 64
        a = C.c_double(0)
        call("ptmf_a11_time", self.handle, prec, x.ptr, y.ptr, int(reps), C.byref(a))
        return a.value

    def time_kernel(self, x: DeviceArray, y: DeviceArray, reps: int) -> float:
        """ms per launch of the element kernel alone (HIP events)."""
        require("MatA11MF.time_kernel", 3 * self.nnodes, x.dtype, x, y)
        prec = 32 if x.dtype == np.float32 else 64
        a = C.c_double(0)
        call("ptmf_a11_time_kernel", self.handle, prec, x.ptr, y.ptr, int(reps), C.byref(a))
        return a.value

    def destroy(self) -> None:
        if getattr(self, "_owned", True) and self.handle is not None and self.handle.value:
            call("ptmf_op_destroy", C.byref(self.handle))
        self.handle = None

    def __del__(self):
        try:
            if (getattr(self, "_owned", True) and self.handle is not None and self.handle.value
                    and _lib._lib is not None):
                _lib._lib.ptmf_op_destroy(C.byref(self.handle))
        except Exception:
            pass


class MatStokesMF(MatA11MF):
    """Adds the P1disc pressure couplings of the Stokes operator."""

    def mult_a12(self, p: DeviceArray, yu: DeviceArray) -> None:
        """MatMult_MFStokes_A12: yu = B p, Dirichlet velocity rows zeroed."""
        require("MatStokesMF.mult_a12", 4 * self.nel, np.float64, p)
        require("MatStokesMF.mult_a12", 3 * self.nnodes, np.float64, yu)
        call("ptmf_a12_mult", self.handle, p.ptr, yu.ptr)

    def mult_a21(self, u: DeviceArray, yp: DeviceArray) -> None:
        """MatMult_MFStokes_A21: yp = B^T mask(u)."""
        require("MatStokesMF.mult_a21", 3 * self.nnodes, np.float64, u)
        require("MatStokesMF.mult_a21", 4 * self.nel, np.float64, yp)
        call("ptmf_a21_mult", self.handle, u.ptr, yp.ptr)

    def mult_stokes(self, x: DeviceArray, y: DeviceArray) -> None:
        """MatMult_MFStokes_A: y = [A11 A12; A21 0] [u | p] in one pass."""
        require("MatStokesMF.mult_stokes", 3 * self.nnodes + 4 * self.nel, np.float64, x, y)
        call("ptmf_stokes_mult", self.handle, x.ptr, y.ptr)

    def stokes_host(self, u: np.ndarray, p: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        x = DeviceArray.from_host(np.concatenate([np.asarray(u, np.float64), np.asarray(p, np.float64)]))
        y = DeviceArray(x.n)
        self.mult_stokes(x, y)
        out = y.download()
        x.free()
        y.free()
        nu = 3 * self.nnodes
        return out[:nu], out[nu:]

    def a12_host(self, p: np.ndarray) -> np.ndarray:
        dp = DeviceArray.from_host(np.asarray(p, dtype=np.float64))
        dy = DeviceArray(3 * self.nnodes)
        self.mult_a12(dp, dy)
        out = dy.download()
        dp.free()
        dy.free()
        return out

    def a21_host(self, u: np.ndarray) -> np.ndarray:
        du = DeviceArray.from_host(np.asarray(u, dtype=np.float64))
        dy = DeviceArray(4 * self.nel)
        self.mult_a21(du, dy)
        out = dy.download()
        du.free()
        dy.free()
        return out

    def schur_blocks_host(self) -> np.ndarray:
        """Element blocks of the scaled viscous pressure mass matrix, [E, 4, 4]
        (MatAssemble_StokesPC_ScaledMassMatrix, stokes_assembly.c:299-450)."""
        d = DeviceArray(16 * self.nel)
        call("ptmf_schur_mass_blocks", self.handle, d.ptr)
        out = d.download().reshape(-1, 4, 4)
        d.free()
        return out


class BatchA11:
    """MatMult_MFStokes_A11 of many independent boxes in one launch per stage
    (set-up MatA11MF operators of any sizes)."""

    def __init__(self, ops):
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[op.handle.value for op in self.ops])
        h = C.c_void_p()
        call("ptmf_batch_create", arr, len(self.ops), C.byref(h))
        self.handle = h

    @staticmethod
    def _ptrs(arrs):
        return (C.c_void_p * len(arrs))(*[a.ptr for a in arrs])

    def _check(self, xs, ys, precision):
        if len(xs) != len(self.ops) or len(ys) != len(self.ops):
            raise _lib.PtmfError(1, "BatchA11", f"{len(xs)}/{len(ys)} arrays for {len(self.ops)} boxes")
        dt = np.float64 if precision == 64 else np.float32
        for op, x, y in zip(self.ops, xs, ys):
            require("BatchA11", 3 * op.nnodes, dt, x, y)

    def mult(self, xs, ys, precision: int = 64) -> None:
        self._check(xs, ys, precision)
        call("ptmf_batch_mult", self.handle, int(precision), self._ptrs(xs), self._ptrs(ys))

    def time_mult(self, xs, ys, reps: int, precision: int = 64) -> float:
        self._check(xs, ys, precision)
        ms = C.c_double()
        call("ptmf_batch_time", self.handle, int(precision), self._ptrs(xs), self._ptrs(ys), int(reps), C.byref(ms))
        return ms.value

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            call("ptmf_batch_destroy", C.byref(self.handle))
        self.handle = None
