"""GPU parity of the quadrature-point power-law viscosity and the device-side
viscosity upload (Picard update) against the CPU oracle, through the C-ABI.

Tolerance (FP64): ||eta_gpu - eta_oracle||_inf / ||eta_oracle||_inf <= 1e-12;
FP32 MatMult after the upload: <= 5e-5 relative to the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptatin3d_amd import fields as F
from ptatin3d_amd._lib import DeviceArray, PtmfError
from ptatin3d_amd.operators import MatA11MF

from conftest import rel_inf

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
TOL32 = 5e-5


def _case(m, seed=11):
    rng = np.random.default_rng(seed)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    E = m[0] * m[1] * m[2]
    u = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    B = rng.uniform(1.0, 10.0, 27 * E)
    return rng, coords, u, B


@pytest.mark.parametrize("m,n,lo,hi", [((5, 4, 3), 3.0, 1e-6, 1e6), ((1, 1, 1), 2.0, 1e-6, 1e6),
                                       ((7, 2, 5), 3.0, 0.5, 2.0), ((4, 4, 4), 1.0, 1e-6, 1e6)])
def test_powerlaw_parity(gpu, m, n, lo, hi):
    rng, coords, u, B = _case(m)
    A = MatA11MF(*m).setup(coords, np.ones(27 * m[0] * m[1] * m[2]), F.dirichlet_all_faces(*m))
    du, dB = DeviceArray.from_host(u), DeviceArray.from_host(B)
    de = DeviceArray(B.size)
    A.powerlaw_viscosity(du, dB, n, lo, hi, de)
    got = de.download()
    ref = O.powerlaw_eta(*m, coords, u, B, n, lo, hi)
    assert rel_inf(got, ref) <= TOL64
    if lo == 0.5:  # the bounds are active for part of the points
        assert np.any(ref == lo) and np.any(ref == hi)
    A.destroy()


def test_zero_velocity_floor(gpu):
    m = (3, 3, 2)
    _, coords, u, B = _case(m)
    A = MatA11MF(*m).setup(coords, np.ones(B.size), F.dirichlet_all_faces(*m))
    du, dB, de = DeviceArray.from_host(np.zeros_like(u)), DeviceArray.from_host(B), DeviceArray(B.size)
    A.powerlaw_viscosity(du, dB, 3.0, 1e-3, 1e40, de)
    assert rel_inf(de.download(), B * 1e-17 ** (-2.0 / 3.0)) <= 1e-14
    A.destroy()


def test_picard_update_device_resident(gpu):
    """eta(u) computed and uploaded on the device; the next MatMult uses it."""
    m = (6, 5, 4)
    rng, coords, u, B = _case(m)
    bc = F.dirichlet_all_faces(*m)
    A = MatA11MF(*m).setup(coords, np.ones(B.size), bc)
    du, dB, de = DeviceArray.from_host(u), DeviceArray.from_host(B), DeviceArray(B.size)
    A.powerlaw_viscosity(du, dB, 3.0, 1e-4, 1e4, de)
    A.set_viscosity_device(de)
    x = rng.uniform(-1, 1, u.size)
    eta = O.powerlaw_eta(*m, coords, u, B, 3.0, 1e-4, 1e4)
    assert rel_inf(A.mult_host(x), O.a11_matmult(*m, coords, eta, bc, x)) <= TOL64
    # the FP32 copy of the viscosity is refreshed by the same upload
    assert rel_inf(A.mult_host(x, precision=32), O.a11_matmult(*m, coords, eta, bc, x)) <= TOL32
    # the partial-assembly variant rebuilds its factors after the state bump
    A.set_kernel_variant(2)
    assert rel_inf(A.mult_host(x), O.a11_matmult(*m, coords, eta, bc, x)) <= TOL64
    A.destroy()


def test_powerlaw_argument_errors(gpu):
    m = (2, 2, 2)
    _, coords, u, B = _case(m)
    A = MatA11MF(*m).setup(coords, np.ones(B.size), F.dirichlet_all_faces(*m))
    du, dB, de = DeviceArray.from_host(u), DeviceArray.from_host(B), DeviceArray(B.size)
    for n, lo, hi in [(0.0, 1e-3, 1e3), (3.0, 0.0, 1e3), (3.0, 10.0, 1.0)]:
        with pytest.raises(PtmfError):
            A.powerlaw_viscosity(du, dB, n, lo, hi, de)
    A.destroy()
