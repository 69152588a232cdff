"""Autotuner for the A11 MatMult work decomposition.

The MatMult kernel assigns a TX x TY tile of elements (TX*TY <= 21, one
64-lane wave, 3 lanes per element) to a wave that marches through ZC element
layers.  The best shape depends on the box (partial tiles waste lanes, thin
tiles raise face-buffer traffic, long z-chunks leave too few waves to fill
256 CUs), and so does the choice between the plain and the software-pipelined
kernel.  `tune(op, ...)` times candidate (shape, variant) pairs with HIP
events on the operator's own stream and installs the fastest one; results are
cached per (mesh, precision) in memory.
"""
from __future__ import annotations

import itertools

import numpy as np

from ._lib import DeviceArray

_CACHE: dict[tuple, tuple[tuple[int, int, int], int]] = {}
VARIANTS = (0, 1)


def candidates(mx: int, my: int, mz: int, max_shapes: int = 8) -> list[tuple[int, int, int]]:
    """Tile shapes ranked by lane fill, each combined with a few z-chunks."""
    shapes = []
    for TX, TY in itertools.product(range(1, 22), range(1, 22)):
        if TX * TY > 21 or (2 * TX + 1) * (2 * TY + 1) > 130 or TX > mx or TY > my:
            continue
        tiles = -(-mx // TX) * -(-my // TY)
        fill = mx * my / (tiles * 21.0)
        shapes.append((fill - 0.02 * (1.0 / TX + 1.0 / TY), TX, TY))
    shapes.sort(reverse=True)
    out = []
    for _, TX, TY in shapes[:max_shapes]:
        for ZC in (1, 2, 4, 8, 16):
            if ZC <= mz:
                out.append((TX, TY, ZC))
    return out


def tune(op, precision: int = 64, reps: int = 10, max_shapes: int = 8, verbose: bool = False,
         variants=VARIANTS):
    """Time the candidate (tiling, kernel variant) pairs on `op` and keep the
    fastest.  Returns ((TX, TY, ZC), variant) and the sorted timings."""
    key = (op.mx, op.my, op.mz, precision)
    if key in _CACHE:
        t, v = _CACHE[key]
        op.set_tiling(*t)
        op.set_kernel_variant(v)
        return _CACHE[key], []
    dt = np.float64 if precision == 64 else np.float32
    n = 3 * op.nnodes
    x = DeviceArray.from_host(np.random.default_rng(0).uniform(-1, 1, n).astype(dt))
    y = DeviceArray(n, dt)
    results = []
    for v in variants:
        op.set_kernel_variant(v)
        for t in candidates(op.mx, op.my, op.mz, max_shapes):
            op.set_tiling(*t)
            op.time_mult(x, y, 2)
            ms = op.time_mult(x, y, reps)
            results.append((ms, (t, v)))
            if verbose:
                print(f"  tiling {t} variant {v}: {ms * 1e3:.1f} us", flush=True)
    results.sort()
    best = results[0][1]
    op.set_tiling(*best[0])
    op.set_kernel_variant(best[1])
    _CACHE[key] = best
    x.free()
    y.free()
    return best, results
