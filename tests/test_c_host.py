"""The C-ABI used from a plain C program (tests/c/abi_host.c), the way a
PETSc adapter in C would use it: the header compiles as C99 and links against
libptmf.so on any machine; on a GPU the program's MatMult and SpMV results
match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from ptatin3d_amd import _lib, fields as F

from conftest import rel_inf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "abi_host")
    libdir = os.path.dirname(_lib.LIBPATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "c", "abi_host.c"), f"-L{libdir}", "-lptmf", f"-Wl,-rpath,{libdir}",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_host_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_c_host_matches_oracle(gpu, tmp_path):
    from oracle import oracle as O

    exe = _build(tmp_path)
    m = (4, 3, 5)
    rng = np.random.default_rng(17)
    coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
    eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-2, 1e2, rng)
    bc = F.dirichlet_all_faces(*m)
    x = rng.uniform(-1, 1, 3 * F.nnodes(*m))
    paths = {k: os.path.join(str(tmp_path), k + ".bin") for k in ("coords", "eta", "bc", "x", "y", "s")}
    coords.astype(np.float64).tofile(paths["coords"])
    eta.astype(np.float64).tofile(paths["eta"])
    bc.astype(np.uint8).tofile(paths["bc"])
    x.astype(np.float64).tofile(paths["x"])
    r = subprocess.run([exe, *map(str, m), paths["coords"], paths["eta"], paths["bc"], paths["x"], paths["y"],
                        paths["s"]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)
    y = np.fromfile(paths["y"], dtype=np.float64)
    s = np.fromfile(paths["s"], dtype=np.float64)
    assert rel_inf(y, O.a11_matmult(*m, coords, eta, bc, x)) <= 1e-12
    assert rel_inf(s, O.a11_tensor(*m, coords, eta, x)) <= 1e-12
