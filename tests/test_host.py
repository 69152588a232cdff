"""Host-side logic that runs without a GPU: the autotuner's candidate set, the
rift workload's boundary data, and the benchmark's one-GPU contract."""
import json
import os
import subprocess
import sys

import numpy as np

from ptatin3d_amd import autotune
from ptatin3d_amd import fields as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_autotune_candidates_respect_the_kernel_limits():
    """Every candidate fits the tile kernel: at most 21 elements per wave,
    a node window of at most 130 nodes, and no tile larger than the box."""
    for m in ((64, 64, 64), (5, 4, 3), (1, 1, 1), (22, 1, 3), (192, 192, 96)):
        c = autotune.candidates(*m)
        assert c
        for TX, TY, ZC in c:
            assert TX * TY <= 21 and (2 * TX + 1) * (2 * TY + 1) <= 130
            assert 1 <= TX <= m[0] and 1 <= TY <= m[1] and 1 <= ZC <= m[2]


def test_rift_boundary_data():
    """Extension on the x-faces (u_x = -1 at x = 0, +1 at x = 2), free slip
    on the y-faces (u_y = 0) and the base (u_z = 0), free top."""
    mx, my, mz = 3, 2, 2
    X, Y, Z = 2 * mx + 1, 2 * my + 1, 2 * mz + 1
    m, g = F.rift_bc(mx, my, mz)
    m = m.reshape(Z, Y, X, 3).astype(bool)
    g = g.reshape(Z, Y, X, 3)
    assert m[..., 0].sum() == 2 * Y * Z and m[:, :, 0, 0].all() and m[:, :, -1, 0].all()
    assert m[..., 1].sum() == 2 * X * Z and m[:, 0, :, 1].all() and m[:, -1, :, 1].all()
    assert m[..., 2].sum() == X * Y and m[0, :, :, 2].all() and not m[-1, :, :, 2].any()
    assert (g[:, :, 0, 0] == -1.0).all() and (g[:, :, -1, 0] == 1.0).all()
    assert np.count_nonzero(g) == 2 * Y * Z and not g[~m].any()
    # the load is gravity only, denser mantle below the crust
    c = F.config5(mx, my, mz)
    f = c["f"].reshape(-1, 3)
    z = c["coords"].reshape(-1, 3)[:, 2]
    assert not f[:, :2].any() and (f[:, 2] < 0).all()
    assert f[z < 0.75, 2].min() < f[z > 0.75, 2].max()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"] and line["n_gpus"] == 2


def _fake_array(n, dtype):
    """A DeviceArray-shaped object that owns no device memory (its pointer is
    never dereferenced: the wrappers must refuse it before any C call)."""
    import ctypes as C

    from ptatin3d_amd._lib import DeviceArray

    class _NoFree(DeviceArray):
        def __del__(self):
            pass

    a = _NoFree.__new__(_NoFree)
    a.dtype = np.dtype(dtype)
    a.n = int(n)
    a.nbytes = a.n * a.dtype.itemsize
    a.ptr = C.c_void_p(0x1000)
    return a


def test_wrappers_refuse_mismatched_device_arrays():
    """Short arrays or mixed precisions never reach the kernels (they would
    index past the allocation on the GPU): PTMF_ERR_ARG from the wrapper."""
    import pytest

    from ptatin3d_amd._lib import PtmfError
    from ptatin3d_amd.multigrid import GMG
    from ptatin3d_amd.operators import MatA11MF, MatStokesMF

    A = MatStokesMF.__new__(MatStokesMF)
    A.mx, A.my, A.mz = 3, 2, 2
    A.nnodes, A.nel, A.handle = F.nnodes(3, 2, 2), 12, None
    n = 3 * A.nnodes
    bad = [
        lambda: A.mult(_fake_array(n, np.float64), _fake_array(n - 3, np.float64)),
        lambda: A.mult(_fake_array(n, np.float64), _fake_array(n, np.float32)),
        lambda: A.mult(_fake_array(n, np.float32), _fake_array(n, np.float64)),
        lambda: A.get_diagonal(_fake_array(n, np.float32)),
        lambda: A.spmv(_fake_array(n + 3, np.float64), _fake_array(n, np.float64)),
        lambda: A.mult_a12(_fake_array(4 * 12 - 1, np.float64), _fake_array(n, np.float64)),
        lambda: A.mult_a21(_fake_array(n, np.float64), _fake_array(4 * 11, np.float64)),
        lambda: A.mult_stokes(_fake_array(n + 48, np.float64), _fake_array(n, np.float64)),
        lambda: A.spmv_host(np.zeros(n), np.zeros(n - 1)),
        lambda: A.set_viscosity_device(_fake_array(27 * 12 - 1, np.float64)),
    ]
    g = GMG.__new__(GMG)
    g.nlevels, g.dims, g.handle = 2, [(2, 1, 1), (4, 2, 2)], None
    g.n = 3 * F.nnodes(4, 2, 2)
    g.nstokes = g.n + 4 * 16
    nc = 3 * F.nnodes(2, 1, 1)
    bad += [
        lambda: g.vcycle(_fake_array(g.n, np.float64), _fake_array(nc, np.float64)),
        lambda: g.level_apply(0, _fake_array(nc, np.float32), _fake_array(nc, np.float64)),
        lambda: g.level_apply(0, _fake_array(g.n, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.level_apply(2, _fake_array(g.n, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.smooth(1, _fake_array(g.n, np.float32), _fake_array(g.n, np.float64), 2, False),
        lambda: g.prolong_add(1, _fake_array(g.n, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.restrict(1, _fake_array(g.n, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.fgmres(_fake_array(g.n, np.float64), _fake_array(g.n, np.float32)),
        lambda: g.fcg(_fake_array(g.n - 1, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.stokes_fgmres(_fake_array(g.n, np.float64), _fake_array(g.n, np.float64)),
        lambda: g.stokes_apply(_fake_array(g.nstokes, np.float64), _fake_array(g.n, np.float64)),
    ]
    for f in bad:
        with pytest.raises(PtmfError) as ei:
            f()
        assert ei.value.code == 1
    assert isinstance(A, MatA11MF)


def test_autotune_cache_persists(tmp_path, monkeypatch):
    """Tuned choices survive the process: written to the JSON cache, read
    back by a fresh module state, keyed by box, precision and candidates."""
    monkeypatch.setenv("PTMF_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_CACHE", {})
    monkeypatch.setattr(autotune, "_LOADED", True)
    k = autotune._key(64, 64, 64, 32, (0, 1, 3), 4)
    autotune._CACHE[k] = ((4, 5, 2), 3, 2)
    autotune._save()
    monkeypatch.setattr(autotune, "_CACHE", {})
    monkeypatch.setattr(autotune, "_LOADED", False)
    autotune._load()
    assert autotune._CACHE[k] == ((4, 5, 2), 3, 2)
    assert autotune._key(64, 64, 64, 64, (0, 1, 3), 4) not in autotune._CACHE
