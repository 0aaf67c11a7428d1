"""Tile-shape autotuner for MarchingCubesSearch.

Three shapes matter on the GPU: the culling brick (64 x B x B corners:
smaller bricks fold programs tighter but cost more planning), how many grid
rows one walk of a brick program evaluates (more rows amortise the walk,
fewer keep register use and latency down) and how many wavefronts share a
brick (fewer give each wave more rows behind one staging of the program,
more put more waves on the brick's latency chain).  ``tune`` times each candidate on the
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
WAVES_PER_BRICK = (1, 2, 4)


def _time_once(L, ctx, hs, deltas, iters, opts) -> float:
    """Wall time of one call: MarchingCubesSearch of one solid, or the batch
    entry point for several."""
    n = len(hs)
    outs = (ctypes.c_void_p * n)()
    t0 = time.perf_counter()
    if n == 1:
        one = ctypes.c_void_p()
        _lib.check(L.mc3d_marching_cubes_search(ctx, hs[0], deltas[0], iters[0], ctypes.byref(opts), ctypes.byref(one)))
        outs[0] = one.value
    else:
        _lib.check(L.mc3d_marching_cubes_search_batch(ctx, (ctypes.c_void_p * n)(*hs), (ctypes.c_double * n)(*deltas),
                                                      (ctypes.c_int32 * n)(*iters), n, ctypes.byref(opts), outs))
    dt = time.perf_counter() - t0
    for i in range(n):
        L.mc3d_mesh_free(outs[i])
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
    waves_per_brick=WAVES_PER_BRICK,
) -> dict:
    """Returns {"brick", "rows_per_walk", "waves_per_brick", "ms", "table"} for the fastest shape
    (best of ``reps`` wall-clock runs per candidate).  ``s`` may be a list of
    solids (with lists of deltas and iters): the shapes of one batched call."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    solids = list(s) if isinstance(s, (list, tuple)) else [s]
    deltas = [float(d) for d in (delta if isinstance(delta, (list, tuple)) else [delta] * len(solids))]
    its = [int(i) for i in (iters if isinstance(iters, (list, tuple)) else [iters] * len(solids))]
    hs = [x._build(b) for x in solids]
    table = []
    try:
        for brick, rpw, wpb in itertools.product(bricks, rows_per_walk, waves_per_brick):
            o = _lib.mc3d_options()
            o.precision, o.cull, o.device_only = precision, 1, 1
            o.brick, o.rows_per_walk, o.waves_per_brick = brick, rpw, wpb
            _time_once(L, ctx, hs, deltas, its, o)  # warm: buffer sizes, code objects
            best = min(_time_once(L, ctx, hs, deltas, its, o) for _ in range(reps))
            table.append({"brick": brick, "rows_per_walk": rpw, "waves_per_brick": wpb, "ms": 1e3 * best})
    finally:
        for h in hs:
            b.release(h)
        b.close()
    win = min(table, key=lambda r: r["ms"])
    return {"brick": win["brick"], "rows_per_walk": win["rows_per_walk"], "waves_per_brick": win["waves_per_brick"],
            "ms": win["ms"], "table": table}
