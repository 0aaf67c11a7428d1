"""Tile-shape autotuner for MarchingCubesSearch.

Two shapes matter on the GPU: the culling brick (64 x B x B corners: smaller
bricks fold programs tighter but cost more planning) and how many grid rows
one walk of a brick program evaluates (more rows amortise the walk, fewer
keep register use and latency down).  ``tune`` times each candidate on the
actual solid and grid (results on the device, nothing copied back) and
returns the fastest; every candidate produces the same triangles.
"""
from __future__ import annotations

import ctypes
import itertools
import time

from . import _lib
from .solids import LibBackend, Solid

BRICKS = (2, 4, 8)
ROWS_PER_WALK = (1, 2, 4)


def _time_once(L, ctx, h, delta, iters, opts) -> float:
    out = ctypes.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.mc3d_marching_cubes_search(ctx, h, delta, iters, ctypes.byref(opts), ctypes.byref(out)))
    dt = time.perf_counter() - t0
    L.mc3d_mesh_free(out)
    return dt


def tune(
    s: Solid,
    delta: float,
    iters: int,
    precision: int = _lib.MC3D_PREC_FP32_GUARD,
    reps: int = 3,
    device: int = 0,
    bricks=BRICKS,
    rows_per_walk=ROWS_PER_WALK,
) -> dict:
    """Returns {"brick", "rows_per_walk", "ms", "table"} for the fastest shape
    (best of ``reps`` wall-clock runs per candidate)."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    table = []
    try:
        for brick, rpw in itertools.product(bricks, rows_per_walk):
            o = _lib.mc3d_options()
            o.precision, o.cull, o.device_only = precision, 1, 1
            o.brick, o.rows_per_walk = brick, rpw
            _time_once(L, ctx, h, delta, iters, o)  # warm: buffer sizes, code objects
            best = min(_time_once(L, ctx, h, delta, iters, o) for _ in range(reps))
            table.append({"brick": brick, "rows_per_walk": rpw, "ms": 1e3 * best})
    finally:
        b.release(h)
        b.close()
    win = min(table, key=lambda r: r["ms"])
    return {"brick": win["brick"], "rows_per_walk": win["rows_per_walk"], "ms": win["ms"], "table": table}
