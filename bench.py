"""Benchmark: FP64 matrix-free A11 MatMult (MatMult_MFStokes_A11) on one MI355X.

Workload: 64^3 Q2 elements (6,440,067 velocity dofs), affine unit cube, sinker
viscosity (32 seeded spheres of radius 0.05-0.12 at 1e4 in a background of 1,
evaluated at the 27 Gauss points per element), random input velocity with
Dirichlet rows on all faces.  One step = one MatMult with vectors resident in
HBM (column masking, element apply, Dirichlet rows).  Prints one JSON line.

The MatMult tiling and kernel variant are picked by the autotuner before the
timed region.  Further objects of the line:
  batch   the batched MatMult of --batch independent boxes (default 256, 0
          skips it) against one MatMult per box;
  solve   the time-to-1e-8 solve of the A11 block on the layered
          contrast-1e6 workload at --solve-m elements per direction (default
          128): flexible CG in FP64 with a 4-level V-cycle (Chebyshev(3),
          exact FP64 coarse factor), FP32 V-cycle against FP64 V-cycle, both
          to a 1e-8 FP64 residual and compared with a 1e-11 FP64 solve, with
          the heuristic's tilings (--solve-tune: every level's tiling from the
          autotuner in its V-cycle's precision; --no-solve skips the leg);
  stokes  the rift saddle point (--stokes MX MY MZ, default 192 192 96;
          --no-stokes skips it): fused MatMult rate, then FP32- and
          FP64-V-cycle FGMRES solves to 1e-8 and their error against a 1e-11
          FP64 solve (--stokes-schur mass|bdb, --stokes-restart,
          --stokes-maxit).
--copy-probe runs one plain device copy of 256 MiB after the timed region
(the reference kernel for HBM byte counters).  batch_solve: the batched
multigrid solve of --batch-solve boxes (default 256; 0 skips it) against
each box solved on its own.

  python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--m 64] [--no-cpu]
                  [--solve-m 128 | --no-solve] [--solve-tune] [--copy-probe] [--batch-solve 256 | --batch-solve 0]
                  [--stokes 192 192 96 | --no-stokes] [--stokes-schur bdb]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_ELEMENT = 17253  # stokes_operators_tensor.c:281-283

# HBM bytes per launch of the FP64 MatMult kernel measured with rocprofv3
# (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from the plain copy kernel
# of --copy-probe), keyed by (elements per direction, tiling, kernel
# variant); see profiles/a11_tile_kernel_fp64_64cube.md.  Reported only for
# a matching configuration.
MEASURED_TRAFFIC = {
    (64, (4, 5, 2), 0): 269.6e6,
    (64, (7, 3, 2), 0): 275.0e6,
    (64, (5, 4, 2), 0): 270.2e6,
    (64, (3, 7, 2), 0): 276.6e6,
}
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--m", type=int, default=64, help="elements per direction")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default tiling")
    ap.add_argument("--tiling", type=int, nargs=3, metavar=("TX", "TY", "ZC"),
                    help="fixed MatMult tiling (no autotuning)")
    ap.add_argument("--no-solve", action="store_true", help="skip the mixed-precision GMG solve leg")
    ap.add_argument("--solve-m", type=int, default=128, help="elements per direction of the solve leg")
    ap.add_argument("--solve-levels", type=int, default=4)
    ap.add_argument("--solve-cheb", type=int, default=3, help="Chebyshev steps per smoothing")
    ap.add_argument("--solve-krylov", choices=("fgmres", "fcg"), default="fcg",
                    help="outer Krylov method of the solve leg (FGMRES(30) stagnates near 0.9 on this workload)")
    ap.add_argument("--solve-restart", type=int, default=30)
    ap.add_argument("--solve-galerkin", type=int, default=0, help="Galerkin coarse levels of the solve leg")
    ap.add_argument("--solve-tune", action="store_true",
                    help="solve leg with every level's tiling from the autotuner (default: the heuristic's)")
    ap.add_argument("--solve-maxit", type=int, default=4000,
                    help="iteration cap of each solve of the solve leg (a few for a profiled run)")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="boxes in the batched-MatMult leg (0: skip)")
    ap.add_argument("--batch-solve", type=int, default=256,
                    help="boxes in the batched GMG solve leg (default 256, the configured workload; 0 skips it)")
    ap.add_argument("--stokes", type=int, nargs=3, metavar=("MX", "MY", "MZ"), default=[192, 192, 96],
                    help="box of the rift saddle-point leg (default 192 192 96)")
    ap.add_argument("--no-stokes", action="store_true", help="skip the rift saddle-point leg")
    ap.add_argument("--stokes-maxit", type=int, default=1200)
    ap.add_argument("--stokes-restart", type=int, default=100, help="FGMRES restart length of the rift leg")
    ap.add_argument("--stokes-schur", choices=("mass", "bdb"), default="bdb",
                    help="Schur approximation of the rift leg: B^T diag(A11)^-1 B element blocks (default) or the "
                         "reference's scaled viscous mass (FGMRES stalls near 3e-7 with it on this workload)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    from ptatin3d_amd import autotune, fields as F
    from ptatin3d_amd._lib import DeviceArray, call, device_count
    from ptatin3d_amd.operators import MatA11MF

    if device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1

    m = args.m
    c = F.config2(m)
    E = m ** 3
    N = F.nnodes(m, m, m)
    ndof = 3 * N
    A = MatA11MF(m, m, m).setup(c["coords"], c["eta"], c["bc"])
    x = DeviceArray.from_host(c["u"])
    y = DeviceArray(ndof)
    if args.tiling:
        A.set_tiling(*args.tiling)
    elif not args.no_tune:
        # the on-the-fly kernels only: they do the reference's per-apply work
        autotune.tune(A, precision=64, reps=10, max_shapes=4, variants=(0, 1, autotune.COLOUR))
    tiling = A.get_tiling()
    variant = A.get_kernel_variant()
    wg_tiles = A.get_tiles_per_group()

    # warmup, then the timed region: K MatMults bracketed by device syncs
    if args.warmup > 0:
        A.time_mult(x, y, args.warmup)
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    ms_event = A.time_mult(x, y, args.steps)
    call("ptmf_device_synchronize")
    t1 = time.perf_counter()
    ms_wall = 1e3 * (t1 - t0) / args.steps
    ms_step = max(ms_wall, ms_event)
    y_gpu = y.download()

    # dominant kernel alone (HIP events on the operator's stream)
    ms_kernel = A.time_kernel(x, y, args.steps)
    flops = FLOPS_PER_ELEMENT * E
    alg_bytes = 8 * (9 * N + 27 * E)
    tflops = flops / (ms_kernel * 1e-3) / 1e12
    gbs = alg_bytes / (ms_kernel * 1e-3) / 1e9

    kname = {1: "a11_tile_pf_kernel<double, true>", 3: "a11_colour_kernel<double>"}.get(
        variant, "a11_tile_kernel<double, true>" if wg_tiles == 1 else f"a11_tile_wg_kernel<double, true, {wg_tiles}>")

    # the FP32 MatMult of the V-cycle levels on the same workload, tuned in
    # its own precision (reported beside; the value is the FP64 MatMult)
    x32 = DeviceArray.from_host(np.asarray(c["u"], dtype=np.float32))
    y32 = DeviceArray(ndof, np.float32)
    if args.tiling:
        pass
    elif not args.no_tune:
        autotune.tune(A, precision=32, reps=10, max_shapes=4, variants=(0, 1, autotune.COLOUR))
    A.time_mult(x32, y32, max(args.warmup, 1))
    ms32 = A.time_mult(x32, y32, args.steps)
    ms32_kernel = A.time_kernel(x32, y32, args.steps)
    A.mult(x32, y32)
    err32 = float(np.abs(y32.download().astype(np.float64) - y_gpu).max() / np.abs(y_gpu).max())
    tflops32 = flops / (ms32_kernel * 1e-3) / 1e12
    v32, w32 = A.get_kernel_variant(), A.get_tiles_per_group()
    fp32 = {
        "tiling": list(A.get_tiling()), "kernel_variant": v32, "tiles_per_group": w32,
        "ms_matmult": round(ms32, 5), "GDoF_s": round(ndof / (ms32 * 1e-3) / 1e9, 3),
        "kernel": {1: "a11_tile_pf_kernel<float, true>", 3: "a11_colour_kernel<float>"}.get(
            v32, "a11_tile_kernel<float, true>" if w32 == 1 else f"a11_tile_wg_kernel<float, true, {w32}>"),
        "kernel_ms": round(ms32_kernel, 5),
        "achieved_TFLOPs": round(tflops32, 3),
        "peak_TFLOPs": 157.3,
        "frac": round(tflops32 / 157.3, 4),
        "frac_of_nonpacked_vector_fp32": round(tflops32 / PEAK_FP64_TFLOPS, 4),
        "note": "packed FP32 instructions are disabled in this build, so the FP32 VALU ceiling is the "
                "non-packed 78.6 TFLOP/s; 157.3 is MI355X's published FP32 vector peak",
        "rel_inf_vs_fp64": err32,
    }
    x32.free()
    y32.free()
    if not args.tiling and not args.no_tune:
        A.set_tiling(*tiling)
        A.set_kernel_variant(variant)
        A.set_tiles_per_group(wg_tiles)

    # diagonal (setup-time companion of the apply in this workload)
    d = DeviceArray(ndof)
    A.get_diagonal(d)  # first call: set-up of the constant tables
    t2 = time.perf_counter()
    for _ in range(5):
        A.get_diagonal(d)
    ms_diag = 1e3 * (time.perf_counter() - t2) / 5
    # the reference's host-array contract (Yu += A u on host arrays, copies
    # over PCIe included): reported beside, never as the value
    uh = np.ascontiguousarray(c["u"])
    Yu = np.zeros(ndof)
    A.spmv_host(uh, Yu)
    t3 = time.perf_counter()
    for _ in range(5):
        A.spmv_host(uh, Yu)
    ms_host = 1e3 * (time.perf_counter() - t3) / 5

    gdofs = ndof / (ms_step * 1e-3) / 1e9
    out = {
        "metric": "MF A11 GDoF/s (% roofline); time to 1e-8 Stokes solve, FP32-GMG vs FP64",
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh, sinker viscosity, random velocity)",
        "config": {
            "workload": f"FP64 A11 MatMult, {m}^3 Q2 elements, affine unit cube, sinker viscosity "
                        "(32 spheres at 1e4), all-face Dirichlet",
            "elements": E,
            "velocity_dofs": ndof,
            "parallelism": "single GPU",
            "tiling": list(tiling),
            "kernel_variant": variant,
            "tiles_per_group": wg_tiles,
        },
        "roofline": {
            "bound": "mfma",
            "peak_kind": "FP64 78.6 TFLOP/s (vector FMA = matrix peak on MI355X; the kernel uses VALU FMA)",
            "achieved": round(tflops, 3),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(tflops / PEAK_FP64_TFLOPS, 4),
            "traffic": MEASURED_TRAFFIC.get((m, tuple(tiling), variant)),
            "kernel": kname,
            "kernel_ms": round(ms_kernel, 5),
            "alg_flops_per_launch": flops,
            "alg_bytes_per_launch": alg_bytes,
            "alg_hbm_GBs": round(gbs, 1),
            "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
        },
        "timing": {"ms_event": round(ms_event, 5), "ms_wall": round(ms_wall, 5),
                   "ms_diagonal": round(ms_diag, 3),
                   "ms_host_array_spmv_pcie": round(ms_host, 3),
                   "host_array_GDoF_s": round(ndof / (ms_host * 1e-3) / 1e9, 3),
                   "diagonal_GBs": round(8 * (6 * F.nnodes(m, m, m) + 27 * m ** 3) / (ms_diag * 1e-3) / 1e9, 1)},
        "fp32_matmult": fp32,
    }

    if not args.no_cpu:
        from oracle import oracle as O

        # compiled here, for this host's CPU (the library in the tree is a
        # portable build)
        native = O.use_native_build()
        threads = O.num_threads()
        t0 = time.perf_counter()
        nrun = 0
        y_cpu = None
        while True:
            y_cpu = O.a11_matmult(m, m, m, c["coords"], c["eta"], c["bc"], c["u"])
            nrun += 1
            el = time.perf_counter() - t0
            if el >= args.cpu_seconds:
                break
        cpu_s = el / nrun
        out["cpu_baseline"] = {
            "value": round(ndof / cpu_s / 1e9, 5),
            "unit": "GDoF/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{nrun} FP64 A11 MatMults of the same {m}^3 workload "
                      f"({el:.1f} s) on {threads} OpenMP threads, {cpu_model()}, "
                      f"{'-march=native on this host' if native else 'portable x86-64-v3 build'}",
        }
        err = float(np.abs(y_gpu - y_cpu).max() / np.abs(y_cpu).max())
        out["parity_rel_inf_vs_oracle"] = err

    if args.copy_probe:
        nb = 256 << 20
        src, dst = DeviceArray(nb // 8), DeviceArray(nb // 8)
        call("ptmf_memcpy_dtod", dst.ptr, src.ptr, nb)
        out["copy_probe_bytes"] = nb
        src.free()
        dst.free()

    A.destroy()
    out["kernels"] = kernels_leg(m)
    if args.batch > 0:
        out["batch"] = batch_leg(args.batch)
    if not args.no_solve:
        out["solve"] = solve_leg(args.solve_m, args.solve_levels, args.solve_cheb, args.solve_krylov,
                                 args.solve_restart, args.solve_galerkin, args.solve_maxit,
                                 tune=args.solve_tune)
    if args.batch_solve > 0:
        out["batch_solve"] = batch_solve_leg(args.batch_solve)
    if args.stokes and not args.no_stokes:
        out["stokes"] = stokes_leg(tuple(args.stokes), args.stokes_maxit, args.stokes_schur, args.stokes_restart)
    print(json.dumps(out), flush=True)
    return 0


def batch_leg(nbox: int) -> dict:
    """MatMult of nbox independent boxes (each side drawn from {4, 8, ..., 24}
    elements, perturbed coordinates, log-uniform viscosity in [1e-3, 1e3],
    seeded), one batched launch per stage vs one MatMult per box."""
    from ptatin3d_amd import fields as F
    from ptatin3d_amd._lib import DeviceArray, call
    from ptatin3d_amd.operators import BatchA11, MatA11MF

    rng = np.random.default_rng(20261202)
    ops, vec = [], {64: ([], []), 32: ([], [])}
    E = ndof = 0
    for _ in range(nbox):
        m = tuple(int(v) for v in rng.choice([4, 8, 12, 16, 20, 24], size=3))
        coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
        eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-3, 1e3, rng)
        ops.append(MatA11MF(*m).setup(coords, eta, F.dirichlet_all_faces(*m)))
        x = rng.uniform(-1, 1, 3 * F.nnodes(*m))
        for prec, dt in ((64, np.float64), (32, np.float32)):
            vec[prec][0].append(DeviceArray.from_host(x.astype(dt)))
            vec[prec][1].append(DeviceArray(x.size, dt))
        E += m[0] * m[1] * m[2]
        ndof += x.size
    B = BatchA11(ops)
    res = {"boxes": nbox, "elements": E, "velocity_dofs": ndof}
    for prec in (64, 32):
        xs, ys = vec[prec]
        B.time_mult(xs, ys, 3, prec)
        ms_b = B.time_mult(xs, ys, 20, prec)
        for A, x, y in zip(ops, xs, ys):
            A.mult(x, y)
        call("ptmf_device_synchronize")
        t0 = time.perf_counter()
        for _ in range(5):
            for A, x, y in zip(ops, xs, ys):
                A.mult(x, y)
        call("ptmf_device_synchronize")
        ms_s = (time.perf_counter() - t0) / 5 * 1e3
        res[f"fp{prec}"] = {"batched_ms": round(ms_b, 4), "batched_GDoFs": round(ndof / ms_b / 1e6, 2),
                            "per_box_ms": round(ms_s, 4), "per_box_GDoFs": round(ndof / ms_s / 1e6, 2)}
    B.destroy()
    for A in ops:
        A.destroy()
    return res


def kernels_leg(m: int) -> dict:
    """The multigrid's other kernels on the m^3 workload (finest two levels of a
    hierarchy down to 8^3): the
    transfers against their algorithmic bytes, prolong-add s (3 N_c + 6 N_f)
    and restrict s (3 N_f + 3 N_c), and one Chebyshev-Jacobi step (one A11
    MatMult plus the vector streams) in FP64 and FP32; wall clock over
    repeated synchronised calls."""
    from ptatin3d_amd import fields as F
    from ptatin3d_amd._lib import DeviceArray, call
    from ptatin3d_amd.multigrid import GMG, default_options

    c = F.config2(m)
    nl = 1
    while (m >> nl) >= 8 and nl < 4:
        nl += 1
    lf = nl - 1
    mg = GMG(m, m, m, nl)
    mg.set_coordinates(c["coords"])
    mg.set_viscosity(lf, c["eta"])
    mg.coarsen_viscosity("qp-arithmetic")
    mg.set_dirichlet_all_faces()
    mg.setup(default_options(smoother_precision=64))
    nf, nc = F.nnodes(m, m, m), F.nnodes(m // 2, m // 2, m // 2)
    xf, xc = DeviceArray(3 * nf), DeviceArray(3 * nc)
    xf.upload(np.random.default_rng(1).uniform(-1, 1, 3 * nf))
    xc.upload(np.random.default_rng(2).uniform(-1, 1, 3 * nc))

    def timed(f, reps=20):
        f()
        call("ptmf_device_synchronize")
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        call("ptmf_device_synchronize")
        return 1e3 * (time.perf_counter() - t0) / reps

    res = {"workload": f"{m}^3 -> {m // 2}^3, FP64 transfers"}
    ms = timed(lambda: mg.prolong_add(lf, xc, xf))
    res["prolong_add"] = {"ms": round(ms, 4), "GBs": round(8 * (3 * nc + 6 * nf) / (ms * 1e-3) / 1e9, 1)}
    ms = timed(lambda: mg.restrict(lf, xf, xc))
    res["restrict"] = {"ms": round(ms, 4), "GBs": round(8 * (3 * nf + 3 * nc) / (ms * 1e-3) / 1e9, 1)}
    for prec, dt in ((64, np.float64), (32, np.float32)):
        b = DeviceArray.from_host(c["u"].astype(dt))
        x = DeviceArray(b.n, dt)
        x.zero()
        ms1 = timed(lambda: mg.smooth(lf, b, x, 1, True), 10)
        ms5 = timed(lambda: mg.smooth(lf, b, x, 6, True), 10)
        res[f"chebyshev_step_fp{prec}_ms"] = round((ms5 - ms1) / 5, 4)
    mg.destroy()
    return res


def batch_solve_leg(nbox: int) -> dict:
    """Velocity-block solves of nbox independent boxes (each side drawn from
    {4, 8, ..., 24} elements, perturbed coordinates, log-uniform viscosity in
    [1e-3, 1e3], random right-hand side, all-face Dirichlet; seeded): 3-level
    V-cycle (FP32 Chebyshev(3)-Jacobi, FP64 dense coarse inverse), FP64
    flexible CG to 1e-8 per box.  Batched: all boxes in lockstep with one
    launch per stage and the V-cycle replayed as one graph; against each box
    solved on its own (its own graph-captured V-cycle), one after another."""
    from ptatin3d_amd import fields as F
    from ptatin3d_amd._lib import DeviceArray, call
    from ptatin3d_amd.multigrid import GMG, BatchGMG, default_options

    rng = np.random.default_rng(20261203)
    mgs, bs, xs = [], [], []
    E = ndof = 0
    t0 = time.perf_counter()
    for _ in range(nbox):
        m = tuple(int(v) for v in rng.choice([4, 8, 12, 16, 20, 24], size=3))
        coords = F.perturb_interior(F.box_coords(*m), *m, 0.1, rng)
        eta = F.eta_loguniform(m[0] * m[1] * m[2], 1e-3, 1e3, rng)
        mg = GMG(*m, 3)
        mg.set_coordinates(coords)
        mg.set_viscosity(2, eta)
        mg.coarsen_viscosity("qp-arithmetic")
        mg.set_dirichlet_all_faces()
        mg.setup(default_options(smoother_precision=32, pre_its=3, post_its=3))
        mgs.append(mg)
        b = F.random_velocity(*m, F.dirichlet_all_faces(*m), rng)
        bs.append(DeviceArray.from_host(b))
        xs.append(DeviceArray(b.size))
        E += m[0] * m[1] * m[2]
        ndof += b.size
    setup_s = time.perf_counter() - t0
    # tile shape of every box's smoothed levels from the autotuner, in the
    # V-cycle's precision (FP32); one tuning per distinct level box, reused
    # by the boxes of the same shape (the batched MatMult runs variant 0)
    from ptatin3d_amd import autotune
    from ptatin3d_amd.operators import MatA11MF
    t0 = time.perf_counter()
    shapes = set()
    for mg in mgs:
        for lvl in (1, 2):
            dims = mg.dims[lvl]
            op = MatA11MF.borrow(mg.level_op_handle(lvl), *dims)
            autotune.tune(op, precision=32, reps=3, max_shapes=3, variants=(0,), wg_tiles=(1,), persist=False)
            shapes.add(tuple(dims))
    tune_s = time.perf_counter() - t0
    bm = BatchGMG(mgs)
    res = {"boxes": nbox, "elements": E, "velocity_dofs": ndof, "levels": 3,
           "smoother": "FP32 Chebyshev(3)-Jacobi, FP64 dense coarse inverse, FP64 flexible CG to 1e-8",
           "setup_seconds_all_boxes": round(setup_s, 2),
           "tiling": f"autotuned per distinct level box in FP32 ({len(shapes)} boxes, {tune_s:.1f} s)"}
    for x in xs:
        x.zero()
    bm.fcg(bs, xs, rtol=1e-8, maxit=200)  # warm-up: captures the batched V-cycle
    for x in xs:
        x.zero()
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    its, rn = bm.fcg(bs, xs, rtol=1e-8, maxit=200)
    call("ptmf_device_synchronize")
    tb = time.perf_counter() - t0
    bnorm = np.array([np.linalg.norm(b.download()) for b in bs])
    xb = [x.download() for x in xs]
    res["batched"] = {"seconds": round(tb, 4), "dofs_per_s": round(ndof / tb, 1),
                      "its_min": int(its.min()), "its_mean": round(float(its.mean()), 2), "its_max": int(its.max()),
                      "max_rel_residual": float((rn / bnorm).max())}
    print(f"# batch solve: {nbox} boxes, {tb:.3f} s, its {its.min()}..{its.max()}", file=sys.stderr, flush=True)
    bm.destroy()
    # each box on its own (graph-captured V-cycle after the first call)
    its1 = []
    for mg, b, x in zip(mgs, bs, xs):
        x.zero()
        mg.fcg(b, x, rtol=1e-8, maxit=200)
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    for mg, b, x in zip(mgs, bs, xs):
        x.zero()
        n, _, _ = mg.fcg(b, x, rtol=1e-8, maxit=200)
        its1.append(n)
    call("ptmf_device_synchronize")
    ts = time.perf_counter() - t0
    diff = max(float(np.linalg.norm(x.download() - y) / max(np.linalg.norm(y), 1e-300)) for x, y in zip(xs, xb))
    res["per_box"] = {"seconds": round(ts, 4), "dofs_per_s": round(ndof / ts, 1),
                      "its_mean": round(float(np.mean(its1)), 2),
                      "max_rel_diff_vs_batched": diff}
    res["speedup_batched_vs_per_box"] = round(ts / tb, 2)
    # error of the batched 1e-8 solutions against each box's FP64 solve
    # continued to 1e-11 (true FP64 residuals, outside the timed regions)
    errs = []
    for mg, b, x, xbat in zip(mgs, bs, xs, xb):
        x.upload(xbat)
        mg.fcg(b, x, rtol=1e-11, maxit=400)
        xt = x.download()
        errs.append(float(np.linalg.norm(xbat - xt) / max(np.linalg.norm(xt), 1e-300)))
    res["rel_err_vs_tight"] = {"max": max(errs), "mean": float(np.mean(errs))}
    for mg in mgs:
        mg.destroy()
    return res


def solve_leg(m: int, levels: int = 4, cheb: int = 3, krylov: str = "fgmres", restart: int = 30,
              galerkin: int = 0, maxit: int = 4000, tune: bool = False) -> dict:
    """Layered viscosity 1e3/1e1/1e-3 with 64 weak ellipsoids (contrast 1e6),
    smooth body force, all-face Dirichlet, on m^3 Q2 elements.  The A11 solve
    of the configured workload: FGMRES(restart) in FP64 (or flexible CG)
    preconditioned by a `levels`-level geometric V-cycle with
    Chebyshev(cheb)-Jacobi smoothing (esteig 0, 0.2, 0, 1.1 of lambda_max of
    D^-1 A, as the reference's opts), lattice transfers, quadrature-point
    local viscosity means on the rediscretised coarse levels, and an exact
    FP64 coarse solve (multifrontal Cholesky of the assembled coarsest level).
    The V-cycle (operators, smoother, transfers) runs in FP32 or in FP64;
    both solves go to ||b - A x|| <= 1e-8 ||b|| (FP64 residual of the FP64
    operator); each solution is compared with a 1e-11 FP64 solve (the FP64
    run continued from its own solution).  With `tune`, every level's tile
    shape, kernel variant and tiles per work-group come from the autotuner,
    timed in the precision the V-cycle applies that level in (set-up work,
    outside the timed solve)."""
    from ptatin3d_amd import autotune, fields as F
    from ptatin3d_amd._lib import DeviceArray, call
    from ptatin3d_amd.multigrid import GMG, default_options, solve_timed

    c = F.config3(m)
    res = {"workload": f"{m}^3 Q2 elements, layered 1e3/1e1/1e-3 + 64 weak ellipsoids (contrast 1e6), "
                       "smooth body force, all-face Dirichlet",
           "velocity_dofs": 3 * F.nnodes(m, m, m), "levels": levels, "galerkin_levels": galerkin,
           "coarse_level": f"{m >> (levels - 1)}^3",
           "krylov": f"FGMRES({restart})" if krylov == "fgmres" else "flexible CG (FP64)",
           "deviation": None if krylov == "fgmres" else
           ("flexible CG instead of FGMRES: with this V-cycle FGMRES(30) stagnates (relative residual 0.90 "
            "after 4000 iterations at 128^3; 7e-3 after 1500 at 64^3) and FGMRES(100) reaches 1e-4 after 1500 at 64^3"),
           "smoother": f"Chebyshev({cheb})-Jacobi on [0.2, 1.1] lambda_max",
           "transfers": "DMDA Q1 lattice interpolation, R = P^T",
           "coarse_solve": "FP64 multifrontal Cholesky of the assembled coarsest level (exact)",
           "level_tiling": ("every level autotuned in its V-cycle's precision; per run, level: [TX, TY, ZC, "
                            "kernel variant, tiles per work-group]" if tune else "the heuristic's")}
    runs = {}
    bnorm = float(np.linalg.norm(c["b"]))
    for tag, prec in (("fp64", 64), ("fp32", 32)):
        mg = GMG(m, m, m, levels)
        mg.set_coordinates(c["coords"])
        mg.set_viscosity(levels - 1, c["eta"])
        mg.coarsen_viscosity("qp-arithmetic")
        mg.set_dirichlet_all_faces()
        t0 = time.perf_counter()
        mg.setup(default_options(smoother_precision=prec, pre_its=cheb, post_its=cheb, galerkin_levels=galerkin,
                                 coarse_factor=2))
        ts = time.perf_counter() - t0
        info = mg.coarse_info()
        tuned, tt = None, 0.0
        if tune:
            t0 = time.perf_counter()
            tuned = autotune.tune_hierarchy(mg, precision=prec)
            tt = time.perf_counter() - t0
        # the V-cycle alone (the part the smoother precision changes)
        db, dv = DeviceArray.from_host(c["b"]), DeviceArray(c["b"].size)
        mg.vcycle(db, dv)
        call("ptmf_device_synchronize")
        t0 = time.perf_counter()
        for _ in range(5):
            mg.vcycle(db, dv)
        call("ptmf_device_synchronize")
        vms = 1e3 * (time.perf_counter() - t0) / 5
        db.free()
        dv.free()
        x, its, rn, hist, secs = solve_timed(mg, c["b"], rtol=1e-8, maxit=maxit, restart=restart, krylov=krylov)
        runs[tag] = x
        res[tag] = {"its": its, "seconds": round(secs, 4), "setup_seconds": round(ts, 3),
                    "rel_residual": float(rn / bnorm), "vcycle_ms": round(vms, 3),
                    "outer_ms_per_it": round(1e3 * secs / max(its, 1) - vms, 3),
                    "coarse_factor_MB": round(info["bytes"] / 1e6, 1),
                    "coarse_factor_ms": round(info["setup_ms"], 1),
                    "tiling": ({f"L{d['level']}": d["tiling"] + [d["variant"], d["tiles_per_group"]] for d in tuned}
                               if tuned else "heuristic"),
                    "tuning_seconds": round(tt, 2)}
        print(f"# solve leg {tag}: {its} its, {secs:.2f} s, rel residual {rn / bnorm:.2e}",
              file=sys.stderr, flush=True)
        if tag == "fp64":
            xt, its_t, rn_t, _, secs_t = solve_timed(mg, c["b"], rtol=1e-11, maxit=maxit, restart=restart,
                                                     krylov=krylov, x0=x)
            runs["tight"] = xt
            res["tight"] = {"its_after_fp64": its_t, "seconds": round(secs_t, 4), "rel_residual": float(rn_t / bnorm)}
            print(f"# solve leg tight: {its_t} more its, {secs_t:.2f} s, rel residual {rn_t / bnorm:.2e}",
                  file=sys.stderr, flush=True)
        mg.destroy()
    ref = runs["tight"]
    for tag in ("fp64", "fp32"):
        res[tag]["rel_err_vs_tight"] = float(np.linalg.norm(runs[tag] - ref) / np.linalg.norm(ref))
    res["speedup_fp32_vs_fp64"] = round(res["fp64"]["seconds"] / res["fp32"]["seconds"], 3)
    return res


def stokes_leg(dims, maxit: int, schur: str = "mass", restart: int = 40) -> dict:
    """Rift saddle point: a 1e3 crust (z > 0.75) over a 1e0 mantle with a
    seeded 1e-3 notch, power-law n = 3 viscosity of a seeded extension
    velocity at the quadrature points (clipped to [1e-3, 1e3]), gravity load;
    free-slip sides and base, extension velocity on the x-faces, free top.
    FGMRES(restart) in FP64 on [u | p] with the fused saddle-point MatMult and the
    upper block-triangular preconditioner (Schur approximation: the scaled
    viscous mass, or element blocks of B^T diag(A11)^-1 B; one 5-level
    V-cycle for A11: Chebyshev(4)-Jacobi, FP32 or FP64).  Reports the fused MatMult rate, then both solves to 1e-8 or maxit."""
    from ptatin3d_amd import fields as F
    from ptatin3d_amd._lib import DeviceArray, call
    from ptatin3d_amd.multigrid import (default_options, powerlaw_viscosity, rift_hierarchy,
                                        stokes_rhs_lifted, stokes_solve_timed)
    from ptatin3d_amd.operators import MatStokesMF

    mx, my, mz = dims
    nl = 1
    while all(((d >> nl) << nl) == d and (d >> nl) >= 2 for d in dims) and nl < 5:
        nl += 1
    c = F.config5(mx, my, mz)
    eta = powerlaw_viscosity(dims, c["coords"], c["u0"], c["B"], c["nexp"], c["eta_min"], c["eta_max"])
    bu, bp = stokes_rhs_lifted(dims, c["coords"], eta, c["bc"], c["g"], c["f"])
    E, nu = mx * my * mz, bu.size
    res = {"workload": f"rift saddle point {mx}x{my}x{mz} Q2-P1disc: 1e3 crust / 1e0 mantle / 1e-3 notch, "
                       "power-law n=3 viscosity of a seeded extension velocity (contrast 1e6), gravity, "
                       "free-slip sides and base, x-face extension, free top",
           "velocity_dofs": nu, "pressure_dofs": 4 * E, "levels": nl,
           "krylov": f"FGMRES({restart}), FP64, fused saddle-point MatMult",
           "preconditioner": "upper block-triangular: " + {"mass": "scaled viscous-mass",
                                                            "bdb": "B^T diag(A11)^-1 B element-block"}[schur]
                             + " Schur (FP64), one V-cycle for A11; with the FP32 V-cycle the A12 z_p product "
                               "that forms the V-cycle's right-hand side is FP32 too"}
    if schur == "bdb":
        res["deviation"] = ("Schur approximation: element blocks of B^T diag(A11)^-1 B (no counterpart in the "
                            "reference) instead of the reference's scaled viscous mass, with which FGMRES stalls "
                            "near 3e-7 on this workload")
    # fused saddle-point MatMult throughput
    A = MatStokesMF(mx, my, mz).setup(c["coords"], eta, c["bc"])
    x = DeviceArray.from_host(np.concatenate([bu, bp]))
    y = DeviceArray(x.n)
    for _ in range(3):
        A.mult_stokes(x, y)
    call("ptmf_device_synchronize")
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        A.mult_stokes(x, y)
    call("ptmf_device_synchronize")
    ms = 1e3 * (time.perf_counter() - t0) / reps
    res["matmult_ms"] = round(ms, 3)
    res["matmult_GDoF_s"] = round((nu + 4 * E) / (ms * 1e-3) / 1e9, 3)
    x.free()
    y.free()
    A.destroy()
    bn = float(np.linalg.norm(np.concatenate([bu, bp])))
    sol = {}
    for tag, prec in (("fp32", 32), ("fp64", 64)):
        t0 = time.perf_counter()
        mg = rift_hierarchy(c, nl, eta, default_options(smoother_precision=prec, pre_its=4, post_its=4,
                                                          esteig=(0.0, 0.1, 0.0, 1.1)), schur=schur)
        ts = time.perf_counter() - t0
        # the V-cycle alone (the part the smoother precision changes)
        db, dv = DeviceArray.from_host(bu), DeviceArray(nu)
        mg.vcycle(db, dv)
        call("ptmf_device_synchronize")
        t0 = time.perf_counter()
        for _ in range(5):
            mg.vcycle(db, dv)
        call("ptmf_device_synchronize")
        vms = 1e3 * (time.perf_counter() - t0) / 5
        db.free()
        dv.free()
        u, p, its, rn, hist, secs = stokes_solve_timed(mg, bu, bp, rtol=1e-8, maxit=maxit, restart=restart)
        if tag == "fp64":
            # the FP64 reference: the same FP64 solve continued from its own
            # solution to a 1e-11 residual
            ut, pt, its_t, rn_t, _, secs_t = stokes_solve_timed(mg, bu, bp, rtol=1e-11, maxit=maxit, restart=restart,
                                                                x0=np.concatenate([u, p]))
            sol["tight"] = (ut, pt)
            res["tight"] = {"its_after_fp64": its_t, "seconds": round(secs_t, 3), "rel_residual": float(rn_t / bn)}
            print(f"# stokes leg tight: {its_t} more its, rel residual {rn_t / bn:.2e}", file=sys.stderr, flush=True)
        mg.destroy()
        sol[tag] = (u, p)
        res[tag] = {"its": its, "seconds": round(secs, 3), "ms_per_it": round(1e3 * secs / max(its, 1), 2),
                    "vcycle_ms": round(vms, 2),
                    "outer_ms_per_it": round(1e3 * secs / max(its, 1) - vms, 2),
                    "setup_seconds": round(ts, 3), "rel_residual": float(rn / bn),
                    "converged": bool(rn <= 1e-8 * bn)}
        print(f"# stokes leg {tag}: {its} its, {secs:.2f} s, rel residual {rn / bn:.2e}",
              file=sys.stderr, flush=True)
    ut, pt = sol["tight"]
    for tag in ("fp32", "fp64"):
        res[tag]["rel_err_vs_tight"] = {
            "u": float(np.linalg.norm(sol[tag][0] - ut) / np.linalg.norm(ut)),
            "p": float(np.linalg.norm(sol[tag][1] - pt) / np.linalg.norm(pt))}
    res["rel_diff_fp32_vs_fp64"] = {
        "u": float(np.linalg.norm(sol["fp32"][0] - sol["fp64"][0]) / np.linalg.norm(sol["fp64"][0])),
        "p": float(np.linalg.norm(sol["fp32"][1] - sol["fp64"][1]) / np.linalg.norm(sol["fp64"][1]))}
    res["speedup_fp32_vs_fp64_per_it"] = round(res["fp64"]["ms_per_it"] / res["fp32"]["ms_per_it"], 3)
    if res["fp32"]["converged"] and res["fp64"]["converged"]:
        res["speedup_fp32_vs_fp64_time_to_1e-8"] = round(res["fp64"]["seconds"] / res["fp32"]["seconds"], 3)
    return res


if __name__ == "__main__":
    sys.exit(main())
