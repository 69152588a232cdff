"""ptatin3d_amd -- MI355X-native matrix-free Q2-P1disc Stokes operators (gfx950).

The compute path is libptmf.so (HIP, built by ``ptatin3d_amd.build``) behind the
C-ABI of include/ptmf.h; this package holds its build script, its ctypes
binding and the host-side mirror of the reference's operator objects.
"""
from .operators import MatA11MF, MatStokesMF  # noqa: F401
from . import fields  # noqa: F401

__all__ = ["MatA11MF", "MatStokesMF", "fields"]
