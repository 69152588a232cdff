"""ctypes binding of the ptmf C-ABI (include/ptmf.h).

The product path is libptmf.so only: if it is missing or has no GPU to run
on, calls raise instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libptmf.so")

PTMF_OK = 0
ERRORS = {
    1: "PTMF_ERR_ARG",
    2: "PTMF_ERR_NOT_SETUP",
    3: "PTMF_ERR_DEVICE",
    4: "PTMF_ERR_NOMEM",
    5: "PTMF_ERR_UNSUPPORTED",
    6: "PTMF_ERR_DIVERGED",
}

# exported symbol -> (restype, argtypes); every one is declared in include/ptmf.h
_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_D = C.c_double
_PD = C.POINTER(C.c_double)
SIGNATURES: dict[str, tuple] = {
    "ptmf_last_error": (C.c_char_p, []),
    "ptmf_version": (_I, [C.POINTER(_I), C.POINTER(_I)]),
    "ptmf_device_count": (_I, [C.POINTER(_I)]),
    "ptmf_set_device": (_I, [_I]),
    "ptmf_malloc": (_I, [C.POINTER(_P), C.c_size_t]),
    "ptmf_free": (_I, [_P]),
    "ptmf_memcpy_htod": (_I, [_P, _P, C.c_size_t]),
    "ptmf_memcpy_dtoh": (_I, [_P, _P, C.c_size_t]),
    "ptmf_memcpy_dtod": (_I, [_P, _P, C.c_size_t]),
    "ptmf_memset": (_I, [_P, _I, C.c_size_t]),
    "ptmf_device_synchronize": (_I, []),
    "ptmf_op_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "ptmf_op_destroy": (_I, [C.POINTER(_P)]),
    "ptmf_op_sizes": (_I, [_P, C.POINTER(_L), C.POINTER(_L)]),
    "ptmf_op_set_coordinates": (_I, [_P, _P]),
    "ptmf_op_set_viscosity": (_I, [_P, _P]),
    "ptmf_op_set_dirichlet": (_I, [_P, _P]),
    "ptmf_op_set_coordinates_device": (_I, [_P, _P]),
    "ptmf_op_set_viscosity_device": (_I, [_P, _P]),
    "ptmf_op_powerlaw_viscosity": (_I, [_P, _P, _P, _D, _D, _D, _P]),
    "ptmf_op_project_points": (_I, [_P, _L, _P, _P, _P, _P]),
    "ptmf_op_state": (_I, [_P, C.POINTER(C.c_ulong)]),
    "ptmf_op_set_tiling": (_I, [_P, _I, _I, _I]),
    "ptmf_op_set_kernel_variant": (_I, [_P, _I]),
    "ptmf_op_get_kernel_variant": (_I, [_P, _P]),
    "ptmf_op_set_tiles_per_group": (_I, [_P, _I]),
    "ptmf_op_get_tiles_per_group": (_I, [_P, C.POINTER(_I)]),
    "ptmf_op_get_tiling": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "ptmf_a11_spmv": (_I, [_P, _P, _P]),
    "ptmf_a11_spmv_host": (_I, [_P, _P, _P]),
    "ptmf_a11_mult": (_I, [_P, _P, _P]),
    "ptmf_a11_mult_f32": (_I, [_P, _P, _P]),
    "ptmf_a11_mult_noalloc": (_I, [_P, _P, _P]),
    "ptmf_op_synchronize": (_I, [_P]),
    "ptmf_a11_diagonal": (_I, [_P, _P]),
    "ptmf_a11_l1_diagonal": (_I, [_P, _P]),
    "ptmf_a12_mult": (_I, [_P, _P, _P]),
    "ptmf_a21_mult": (_I, [_P, _P, _P]),
    "ptmf_stokes_mult": (_I, [_P, _P, _P]),
    "ptmf_a11_time": (_I, [_P, _I, _P, _P, _I, _PD]),
    "ptmf_mg_create": (_I, [_I, _I, _I, _I, C.POINTER(_P)]),
    "ptmf_mg_destroy": (_I, [C.POINTER(_P)]),
    "ptmf_mg_default_options": (_I, [_P]),
    "ptmf_mg_levels": (_I, [_P, C.POINTER(_I)]),
    "ptmf_mg_level_op": (_I, [_P, _I, C.POINTER(_P)]),
    "ptmf_mg_set_coordinates": (_I, [_P, _P]),
    "ptmf_mg_set_viscosity": (_I, [_P, _I, _P]),
    "ptmf_mg_coarsen_viscosity": (_I, [_P, _I]),
    "ptmf_mg_set_dirichlet": (_I, [_P, _I, _P]),
    "ptmf_mg_setup": (_I, [_P, _P]),
    "ptmf_mg_coarse_info": (_I, [_P, C.POINTER(_I), C.POINTER(_L), C.POINTER(_I), _PD]),
    "ptmf_mg_level_eigs": (_I, [_P, _I, _PD, _PD]),
    "ptmf_mg_set_level_eigs": (_I, [_P, _I, _D, _D]),
    "ptmf_mg_vcycle": (_I, [_P, _P, _P]),
    "ptmf_mg_level_apply": (_I, [_P, _I, _I, _P, _P]),
    "ptmf_mg_level_assemble": (_I, [_P, _I, C.POINTER(_L), _P, _P, _P]),
    "ptmf_mg_level_rigid_body_modes": (_I, [_P, _I, _P]),
    "ptmf_mg_prolong_add": (_I, [_P, _I, _P, _P]),
    "ptmf_mg_restrict": (_I, [_P, _I, _P, _P]),
    "ptmf_mg_smooth": (_I, [_P, _I, _I, _I, _I, _P, _P]),
    "ptmf_mg_fgmres": (_I, [_P, _P, _P, _D, _I, _I, C.POINTER(_I), _PD, _PD, _I]),
    "ptmf_mg_fcg": (_I, [_P, _P, _P, _D, _I, C.POINTER(_I), _PD, _PD, _I]),
    "ptmf_batch_create": (_I, [_P, _I, _P]),
    "ptmf_batch_destroy": (_I, [_P]),
    "ptmf_batch_mult": (_I, [_P, _I, _P, _P]),
    "ptmf_batch_time": (_I, [_P, _I, _P, _P, _I, _PD]),
    "ptmf_bmg_create": (_I, [_P, _I, _P]),
    "ptmf_bmg_destroy": (_I, [_P]),
    "ptmf_bmg_vcycle": (_I, [_P, _P, _P]),
    "ptmf_bmg_fcg": (_I, [_P, _P, _P, _D, _I, _P, _P]),
    "ptmf_mg_stokes_setup": (_I, [_P, _I, _D]),
    "ptmf_mg_stokes_set_schur": (_I, [_P, _I]),
    "ptmf_mg_stokes_apply": (_I, [_P, _P, _P]),
    "ptmf_mg_stokes_fgmres": (_I, [_P, _P, _P, _D, _I, _I, C.POINTER(_I), _PD, _PD, _I]),
    "ptmf_schur_mass_blocks": (_I, [_P, _P]),
    "ptmf_a11_time_kernel": (_I, [_P, _I, _P, _P, _I, _PD]),
}


class MGOptions(C.Structure):
    """ptmf_mg_options (include/ptmf.h)."""

    _fields_ = [
        ("levels_mg", C.c_int),
        ("smoother_precision", C.c_int),
        ("pre_its", C.c_int),
        ("post_its", C.c_int),
        ("esteig", C.c_double * 4),
        ("esteig_its", C.c_int),
        ("coarse_rtol", C.c_double),
        ("coarse_maxit", C.c_int),
        ("galerkin_levels", C.c_int),
        ("bc_transfer", C.c_int),
        ("smoother_diag", C.c_int),
        ("use_graph", C.c_int),
        ("transfer", C.c_int),
        ("coarse_factor", C.c_int),
    ]


class PtmfError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libptmf.so (built by ptatin3d_amd.build); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            f"{LIBPATH} is missing: build it with `python -m ptatin3d_amd.build` "
            "(there is no fallback implementation)"
        )
    lib = C.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != PTMF_OK:
        msg = load().ptmf_last_error().decode(errors="replace")
        raise PtmfError(code, where, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def device_count() -> int:
    n = C.c_int(0)
    call("ptmf_device_count", C.byref(n))
    return n.value


class DeviceArray:
    """A device allocation made through the library (ptmf_malloc)."""

    def __init__(self, n: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        call("ptmf_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        call("ptmf_memcpy_htod", self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        call("ptmf_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes))
        return out

    def zero(self) -> None:
        call("ptmf_memset", self.ptr, 0, C.c_size_t(self.nbytes))

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            call("ptmf_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr is not None and self.ptr.value and _lib is not None:
                _lib.ptmf_free(self.ptr)
        except Exception:
            pass


def require(where: str, n: int, dtype, *arrays) -> None:
    """Refuse device arrays that the C function would index past their end or
    read in the wrong precision: each must be a live DeviceArray of exactly n
    elements of dtype (PTMF_ERR_ARG, before anything reaches the GPU)."""
    dt = np.dtype(dtype)
    for a in arrays:
        if not isinstance(a, DeviceArray) or a.ptr is None or not a.ptr.value:
            raise PtmfError(1, where, "expected a live DeviceArray")
        if a.dtype != dt:
            raise PtmfError(1, where, f"array of {a.dtype}, expected {dt}")
        if a.n != int(n):
            raise PtmfError(1, where, f"array of {a.n} elements, expected {int(n)}")


def _hp(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)
