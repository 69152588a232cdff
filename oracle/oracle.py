"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBPATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = C.CDLL(LIBPATH)
        _lib.orc_num_threads.restype = C.c_int
    return _lib


def use_native_build() -> bool:
    """Compile the oracle for this host's CPU (-march=native) and make every
    later call use it; for the CPU-baseline timing, on the host that times
    it.  Returns False (portable build kept) when no compiler is available.
    Call before any other oracle function."""
    global _lib
    native = os.path.join(HERE, "liboracle_native.so")
    r = subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(native):
        return False
    _lib = C.CDLL(native)
    _lib.orc_num_threads.restype = C.c_int
    return True


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def num_threads() -> int:
    return load().orc_num_threads()


def set_num_threads(n: int) -> None:
    load().orc_set_num_threads(C.c_int(n))


def nn(mx, my, mz):
    return (2 * mx + 1) * (2 * my + 1) * (2 * mz + 1)


def a11_tensor(mx, my, mz, coords, eta, u, y=None):
    """y += A11 u (no BCs), tensor path."""
    y = np.zeros(3 * nn(mx, my, mz)) if y is None else y
    load().orc_a11_tensor(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_f64(u)), _p(y))
    return y


def a11_ref(mx, my, mz, coords, eta, u, y=None):
    y = np.zeros(3 * nn(mx, my, mz)) if y is None else y
    load().orc_a11_ref(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_f64(u)), _p(y))
    return y


def a11_element_matrix(mx, my, mz, coords, eta, e):
    Ae = np.zeros((81, 81))
    load().orc_a11_element_matrix(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), C.c_long(e), _p(Ae))
    return Ae


def a11_asm_matmult(mx, my, mz, coords, eta, bc, x):
    y = np.zeros(3 * nn(mx, my, mz))
    load().orc_a11_asm_matmult(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_u8(bc)), _p(_f64(x)), _p(y))
    return y


def a11_matmult(mx, my, mz, coords, eta, bc, x):
    y = np.zeros(3 * nn(mx, my, mz))
    load().orc_a11_matmult(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_u8(bc)), _p(_f64(x)), _p(y))
    return y


def a11_diagonal(mx, my, mz, coords, eta, bc):
    d = np.zeros(3 * nn(mx, my, mz))
    load().orc_a11_diagonal(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_u8(bc)), _p(d))
    return d


def a12_matmult(mx, my, mz, coords, bc, p):
    y = np.zeros(3 * nn(mx, my, mz))
    load().orc_a12_matmult(mx, my, mz, _p(_f64(coords)), _p(_u8(bc)), _p(_f64(p)), _p(y))
    return y


def a21_matmult(mx, my, mz, coords, bc, u):
    y = np.zeros(4 * mx * my * mz)
    load().orc_a21_matmult(mx, my, mz, _p(_f64(coords)), _p(_u8(bc)), _p(_f64(u)), _p(y))
    return y


def schur_mass_blocks(mx, my, mz, coords, eta):
    M = np.zeros(16 * mx * my * mz)
    load().orc_schur_mass_blocks(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(M))
    return M.reshape(-1, 4, 4)


def schur_bdb_blocks(mx, my, mz, coords, eta, bc):
    """Inverse element blocks (-B_e^T D^-1 B_e)^-1 of the Schur approximation
    built from the pressure-gradient columns B = A12 (orc_a12_matmult) and
    the A11 diagonal D (orc_a11_diagonal), Dirichlet rows left out.  The P1
    pressure is discontinuous, so column 4e+i of the assembled A12 holds
    element e's contribution only; the columns of all elements of one parity
    class (ei, ej, ek mod 2) touch disjoint nodes and come from one product."""
    X, Y = 2 * mx + 1, 2 * my + 1
    E = mx * my * mz
    d = a11_diagonal(mx, my, mz, coords, eta, bc)
    free = ~np.asarray(bc).astype(bool)
    l = np.arange(3)
    loc = (l[None, None, :] + X * l[None, :, None] + X * Y * l[:, None, None]).reshape(-1)
    S = np.zeros((E, 4, 4))
    cols = np.zeros((E, 81, 4))
    e = np.arange(E)
    ei, ej, ek = e % mx, (e // mx) % my, e // (mx * my)
    nodes = (2 * ei + X * 2 * ej + X * Y * 2 * ek)[:, None] + loc[None, :]
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(E, 81)
    for cls in range(8):
        sel = (ei % 2 == cls & 1) & (ej % 2 == (cls >> 1) & 1) & (ek % 2 == (cls >> 2) & 1)
        for i in range(4):
            p = np.zeros(4 * E)
            p[4 * e[sel] + i] = 1.0
            y = a12_matmult(mx, my, mz, coords, bc, p)
            cols[sel, :, i] = y[dofs[sel]]
    w = np.where(free[dofs], 1.0 / np.where(free, d, 1.0)[dofs], 0.0)
    S = np.einsum("eri,er,erj->eij", cols, w, cols)
    return np.linalg.inv(-S)


def prolong_add(mxc, myc, mzc, xc, yf):
    load().orc_prolong_add(mxc, myc, mzc, _p(_f64(xc)), _p(yf))
    return yf


def restrict_add(mxc, myc, mzc, xf, yc):
    load().orc_restrict_add(mxc, myc, mzc, _p(_f64(xf)), _p(yc))
    return yc


def inject_coords(mxc, myc, mzc, xf):
    xc = np.zeros(3 * nn(mxc, myc, mzc))
    load().orc_inject_coords(mxc, myc, mzc, _p(_f64(xf)), _p(xc))
    return xc


def cheb_jacobi(mx, my, mz, coords, eta, bc, diag, emin, emax, its, b, x0=None):
    x = np.zeros(3 * nn(mx, my, mz)) if x0 is None else np.array(x0, dtype=np.float64)
    load().orc_cheb_jacobi(mx, my, mz, _p(_f64(coords)), _p(_f64(eta)), _p(_u8(bc)), _p(_f64(diag)),
                           C.c_double(emin), C.c_double(emax), C.c_int(its),
                           C.c_int(0 if x0 is None else 1), _p(_f64(b)), _p(x))
    return x


def powerlaw_eta(mx, my, mz, coords, u, B, nexp, eta_min, eta_max):
    """Strain-rate-dependent quadrature-point viscosity, element-major [27E]."""
    eta = np.zeros(27 * mx * my * mz)
    load().orc_powerlaw_eta(mx, my, mz, _p(_f64(coords)), _p(_f64(u)), _p(_f64(B)), C.c_double(nexp),
                            C.c_double(eta_min), C.c_double(eta_max), _p(eta))
    return eta


def q1_nodal_from_qp(mx, my, mz, eta):
    """Fine nodal field of the coefficient hierarchy from element-major qp values."""
    out = np.zeros(nn(mx, my, mz))
    load().orc_q1_nodal_from_qp(mx, my, mz, _p(_f64(eta)), _p(out))
    return out


def q1_restrict_scaled(mxc, myc, mzc, nf, inject=False):
    out = np.zeros(nn(mxc, myc, mzc))
    load().orc_q1_restrict_scaled(mxc, myc, mzc, int(bool(inject)), _p(_f64(nf)), _p(out))
    return out


def q1_to_qp(mx, my, mz, nodal):
    out = np.zeros(27 * mx * my * mz)
    load().orc_q1_to_qp(mx, my, mz, _p(_f64(nodal)), _p(out))
    return out


def q1_projection_hierarchy(dims, eta_fine, inject=False):
    """Per-level element-major qp viscosity (level 0 coarsest) as the
    reference's coefficient hierarchy builds it from fine qp values; the
    finest level keeps its own values."""
    L = len(dims)
    out = [None] * L
    out[L - 1] = np.asarray(eta_fine, dtype=np.float64).copy()
    nodal = q1_nodal_from_qp(*dims[L - 1], eta_fine)
    for l in range(L - 1, 0, -1):
        nodal = q1_restrict_scaled(*dims[l - 1], nodal, inject)
        out[l - 1] = q1_to_qp(*dims[l - 1], nodal)
    return out


def q1_nodal_from_points(mx, my, mz, elem, xi, val):
    """Fine nodal field from material points (element, reference coordinates, value)."""
    out = np.zeros(nn(mx, my, mz))
    el = np.ascontiguousarray(elem, dtype=np.int64)
    load().orc_q1_nodal_from_points(mx, my, mz, C.c_long(el.size), el.ctypes.data_as(C.c_void_p),
                                    _p(_f64(xi)), _p(_f64(val)), _p(out))
    return out
