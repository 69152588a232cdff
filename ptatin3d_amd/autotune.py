"""Autotuner for the A11 MatMult work decomposition.

The MatMult kernel assigns a TX x TY tile of elements (TX*TY <= 21, one
64-lane wave, 3 lanes per element) to a wave that marches through ZC element
layers; the elements of a tile share an LDS node-plane window for the
scatter.  The best shape depends on the box (partial tiles waste lanes, thin
tiles raise face-buffer traffic, long z-chunks leave too few waves to fill
256 CUs), and so do the kernel variant (plain or software-pipelined z-march)
the scatter scheme (the tile window with face buffers, or the colouring of
the elements into 8 parity classes with direct read-add-write sums, variant
3, no tiles) and the number of tiles per work-group (1, 2 or 4 waves, so
TX*TY, 2 TX*TY or 4 TX*TY elements per work-group).  `tune(op, ...)` times
the candidate (shape, variant, tiles per work-group) triples with HIP events on the operator's own stream, in the precision the
operator will be applied in, and installs the fastest; `tune_hierarchy`
does that for every level of a multigrid hierarchy in the V-cycle's
precision.  Choices are cached per (box, precision, candidate set) in memory
and in a JSON file (PTMF_AUTOTUNE_CACHE, default
~/.cache/ptmf/autotune.json), so later processes reuse them.
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np

from ._lib import DeviceArray

_CACHE: dict[str, tuple[tuple[int, int, int], int, int]] = {}
_LOADED = False
VARIANTS = (0, 1)
COLOUR = 3  # the coloured-scatter variant (tiling does not apply)
WG_TILES = (1, 2, 4)  # tiles (waves) per work-group tried with variant 0


def cache_path() -> str:
    return os.environ.get("PTMF_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "ptmf", "autotune.json"))


def _key(mx, my, mz, precision, variants, max_shapes) -> str:
    return f"{mx}x{my}x{mz}:fp{precision}:v{'-'.join(map(str, variants))}:s{max_shapes}"


def _load() -> None:
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    try:
        with open(cache_path()) as f:
            for k, val in json.load(f).items():
                t, v = val[0], val[1]
                _CACHE.setdefault(k, (tuple(t), int(v), int(val[2]) if len(val) > 2 else 1))
    except (OSError, ValueError, TypeError):
        pass


def _save() -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump({k: [list(t), v, w] for k, (t, v, w) in _CACHE.items()}, f, indent=0, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs re-tuning


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
         variants=VARIANTS, persist: bool = True, wg_tiles=WG_TILES, refine: int = 3):
    """Time the candidate (tiling, kernel variant, tiles per work-group)
    triples on `op` in the given precision and keep the fastest.  Variant 3
    (coloured scatter) is one candidate whatever the tiling; tiles per
    work-group apply to variant 0.  The `refine` fastest of the first pass are
    timed again (best of three runs of 2*reps MatMults each), so that
    candidates a few per cent apart are not ranked by one noisy sample.
    Returns ((TX, TY, ZC), variant, tiles) and the sorted timings (empty when
    the choice came from the cache)."""
    _load()
    key = _key(op.mx, op.my, op.mz, precision, variants, max_shapes) + f":w{'-'.join(map(str, wg_tiles))}"
    if key in _CACHE:
        t, v, w = _CACHE[key]
        op.set_tiling(*t)
        op.set_kernel_variant(v)
        op.set_tiles_per_group(w)
        return _CACHE[key], []
    dt = np.float64 if precision == 64 else np.float32
    n = 3 * op.nnodes
    x = DeviceArray.from_host(np.random.default_rng(0).uniform(-1, 1, n).astype(dt))
    y = DeviceArray(n, dt)
    results = []
    base_tiling = op.get_tiling()
    cands = []
    for v in variants:
        if v == COLOUR:
            cands.append((tuple(base_tiling), v, 1))
            continue
        for t in candidates(op.mx, op.my, op.mz, max_shapes):
            for w in (wg_tiles if v == 0 else (1,)):
                cands.append((tuple(t), v, w))
    for t, v, w in cands:
        op.set_kernel_variant(v)
        op.set_tiling(*t)
        op.set_tiles_per_group(w)
        op.time_mult(x, y, 2)
        ms = op.time_mult(x, y, reps)
        results.append((ms, (t, v, w)))
        if verbose:
            print(f"  tiling {t} variant {v} tiles/group {w}: {ms * 1e3:.1f} us", flush=True)
    results.sort()
    if refine > 1 and len(results) > 1:
        again = []
        for _, (t, v, w) in results[:refine]:
            op.set_kernel_variant(v)
            op.set_tiling(*t)
            op.set_tiles_per_group(w)
            op.time_mult(x, y, 2)
            again.append((min(op.time_mult(x, y, 2 * reps) for _ in range(3)), (t, v, w)))
        again.sort()
        results = again + results[refine:]
        if verbose:
            for ms, (t, v, w) in again:
                print(f"  refined: tiling {t} variant {v} tiles/group {w}: {ms * 1e3:.1f} us", flush=True)
    best = results[0][1]
    op.set_tiling(*best[0])
    op.set_kernel_variant(best[1])
    op.set_tiles_per_group(best[2])
    _CACHE[key] = best
    if persist:
        _save()
    x.free()
    y.free()
    return best, results


def tune_hierarchy(mg, precision: int = 32, reps: int = 5, max_shapes: int = 4, variants=VARIANTS + (COLOUR,),
                   verbose: bool = False) -> list[dict]:
    """Tune every rediscretised level operator of a multigrid hierarchy (a
    `GMG`) in the precision its V-cycle applies it in; returns the choice per
    level.  A changed level operator makes the hierarchy drop its captured
    V-cycle graph and capture it again."""
    from .operators import MatA11MF

    out = []
    gal = getattr(getattr(mg, "opt", None), "galerkin_levels", 0)
    for lvl in range(gal, mg.nlevels):  # Galerkin levels apply a stencil, not the tile kernel
        mx, my, mz = mg.dims[lvl]
        op = MatA11MF.borrow(mg.level_op_handle(lvl), mx, my, mz)
        (t, v, w), res = tune(op, precision=precision, reps=reps, max_shapes=max_shapes, variants=variants)
        out.append({"level": lvl, "box": [mx, my, mz], "precision": precision, "tiling": list(t), "variant": v,
                    "tiles_per_group": w, "us": round(res[0][0] * 1e3, 2) if res else None})
        if verbose:
            print(f"level {lvl} {mx}x{my}x{mz}: tiling {t} variant {v} tiles/group {w}", flush=True)
    return out
