"""Benchmark: MarchingCubesSearch corner evals/s + tris/s on one MI355X.

A step is one MarchingCubesSearch (model3d/mc.go:45-49) of the workload on
the GPU, results left in HBM (device_only); the solid is lowered and uploaded
inside the step.  Default workload: the seeded 64-primitive CSG tree clipped
to [-1, 1]^3 at delta 0.005 (402^3 ~ 65M corners), iters 8.

Prints one JSON line (see DESIGN.md for every field).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MarchingCubesSearch corner evals/s + tris/s, FP32+guard vs FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def workload(name):
    from model3d_amd import workloads

    if name == "csg64":
        c = workloads.CSG64_CONFIG
        return workloads.csg64_solid(c["seed"], c["n_prims"]), c
    if name == "sphere":
        return workloads.sphere_solid(), workloads.SPHERE_CONFIG
    if name == "collider":
        c = dict(workloads.COLLIDER_CONFIG)
        solid, delta, ntri = workloads.collider_solid(c["seed"], c["n_prims"], c["mesh_delta"])
        c["delta"], c["mesh_triangles"] = delta, ntri
        return solid, c
    raise SystemExit(f"unknown workload {name}")


def run_gpu(solid, cfg, precision, cull, steps, warmup, count_mismatch=False, tile=None):
    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    L = _lib.lib()
    ctx = _lib.context(0)
    b = LibBackend()
    h = solid._build(b)
    opts = _lib.mc3d_options()
    opts.precision = precision
    opts.cull = 1 if cull else 0
    if tile:
        opts.brick, opts.rows_per_walk = tile["brick"], tile["rows_per_walk"]
    opts.device_only = 1
    opts.count_mismatch = 1 if count_mismatch else 0
    stats = []

    def step(record):
        out = ctypes.c_void_p()
        _lib.check(
            L.mc3d_marching_cubes_search(ctx, h, cfg["delta"], cfg["iters"], ctypes.byref(opts), ctypes.byref(out)),
            "mc3d_marching_cubes_search",
        )
        if record:
            st = _lib.mc3d_stats()
            L.mc3d_mesh_stats(out, ctypes.byref(st))
            stats.append(st.as_dict())
        L.mc3d_mesh_free(out)

    for _ in range(warmup):
        step(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    wall = time.perf_counter() - t0
    b.release(h)
    b.close()
    return wall, stats


def mean(stats, key):
    return sum(s[key] for s in stats) / max(1, len(stats))


def cpu_baseline(solid, cfg):
    from oracle.binding import oracle_search

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    if cfg.get("mesh_triangles"):
        # Brute-force ray parity over the whole mesh is far too slow for the
        # full grid on the CPU: time Contains on a sample of grid corners.
        import numpy as np

        from oracle.binding import oracle_contains

        mn, mx = solid.Min(), solid.Max()
        rng = np.random.default_rng(0)
        pts = rng.uniform(mn, mx, size=(2000, 3))
        t0 = time.perf_counter()
        oracle_contains(solid, pts)
        wall = time.perf_counter() - t0
        return {
            "value": len(pts) / wall,
            "unit": "corner evals/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{len(pts)} Contains calls at random points in the bounds, one thread, brute-force ray "
            f"parity over {cfg['mesh_triangles']} triangles, {wall:.2f} s",
        }
    t0 = time.perf_counter()
    tris, info = oracle_search(solid, cfg["delta"], cfg["iters"], threads=threads)
    wall = time.perf_counter() - t0
    return {
        "value": info["corners"] / wall,
        "unit": "corner evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the full workload once ({info['corners']} corners, {len(tris)} triangles), C restatement of "
        f"mc.go with its plane-pipeline/parallel-search structure, {wall:.2f} s",
        "tris_per_s": len(tris) / wall,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="csg64", choices=["csg64", "sphere", "collider"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU only: this build runs on single-GPU boxes and has no multi-GPU path "
            "(z-slab / per-solid sharding is a design, see DESIGN.md)",
            file=sys.stderr,
        )
        raise SystemExit(2)

    from model3d_amd import _lib

    solid, cfg = workload(args.workload)
    # Tile shapes from the autotuner (untimed), then the timed steps.
    from model3d_amd import autotune

    tuned = {}
    for label, prec in (("fp32_guard", _lib.MC3D_PREC_FP32_GUARD), ("fp64", _lib.MC3D_PREC_FP64)):
        tuned[label] = autotune.tune(solid, cfg["delta"], cfg["iters"], precision=prec)
    best = "fp32_guard" if tuned["fp32_guard"]["ms"] <= tuned["fp64"]["ms"] else "fp64"
    precision = _lib.MC3D_PREC_FP32_GUARD if best == "fp32_guard" else _lib.MC3D_PREC_FP64
    tile = tuned[best]
    wall, st = run_gpu(solid, cfg, precision, True, args.steps, args.warmup, tile=tile)
    corners, tris, verts = st[0]["corners"], st[0]["triangles"], st[0]["vertices"]
    ms = 1e3 * wall / args.steps
    result = {
        "metric": METRIC,
        "value": corners / (wall / args.steps),
        "unit": "corner evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if best == "fp32_guard" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{cfg['name']}_iters{cfg['iters']}",
            "corners": corners,
            "vertices": verts,
            "triangles": tris,
            "bisection_evals": st[0]["bisection_evals"],
            "precision": best,
            "cull": True,
            "brick": tile["brick"],
            "rows_per_walk": tile["rows_per_walk"],
        },
        "tris_per_s": tris / (wall / args.steps),
        "autotune": {
            k: {"brick": v["brick"], "rows_per_walk": v["rows_per_walk"], "best_ms": v["ms"]}
            for k, v in tuned.items()
        },
    }
    stage_keys = ("ms_plan", "ms_corners", "ms_classify", "ms_emit", "ms_bisect", "ms_device")
    stages = {k: mean(st, k) for k in stage_keys}
    result["stages_ms"] = stages
    result["bricks"] = {"total": st[0]["bricks_total"], "active": st[0]["bricks_active"]}

    # Guard statistics from one extra (untimed) run that also evaluates every
    # float32 decision in float64.
    _, sst = run_gpu(solid, cfg, _lib.MC3D_PREC_FP32_GUARD, True, 1, 0, count_mismatch=True)
    g = sst[0]
    result["guard"] = {
        "fp32_atom_evals": g["guard_evals"],
        "fp64_rechecks": g["guard_rechecks"],
        "fp64_fallback_fraction": g["guard_rechecks"] / g["guard_evals"] if g["guard_evals"] else None,
        "fp32_sign_mismatch_rate": g["guard_mismatches"] / g["guard_evals"] if g["guard_evals"] else None,
        "band_violations": g["guard_violations"],
    }

    if not args.no_fp64:
        n64 = max(1, args.steps // 2)
        for label, cull in (("fp64_plain", False), ("fp64_cull", True)):
            wall64, st64 = run_gpu(solid, cfg, _lib.MC3D_PREC_FP64, cull, n64, 1)
            ms64 = 1e3 * wall64 / n64
            result[label] = {
                "corner_evals_per_s": corners / (ms64 / 1e3),
                "tris_per_s": tris / (ms64 / 1e3),
                "ms_per_step": ms64,
                "stages_ms": {k: mean(st64, k) for k in stage_keys},
            }
        result["speedup_vs_fp64_plain"] = result["fp64_plain"]["ms_per_step"] / ms

    # Roofline of the dominant stage (HIP events on the library's stream).
    dom = max(("ms_plan", "ms_corners", "ms_classify", "ms_emit", "ms_bisect"), key=lambda k: stages[k])
    result["roofline"] = roofline(dom, stages[dom], st[0])

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(solid, cfg)
    print(json.dumps(result), flush=True)


def roofline(stage, ms, st):
    """Algorithmic bytes of the dominant stage per launch / its duration.

    corners : writes the occupancy grid, 1 bit per corner  -> N/8 bytes
    classify: reads the grid ~4x (row + 3 neighbour rows), writes 8 B counts
              per word, scans them                          -> ~N/2 + 4*N/64*8
    emit    : reads grid + offsets, writes 8 B per vertex key and 12 B per
              triangle                                     -> N/8*4 + 8V + 12T
    bisect  : reads 8 B key, writes 24 B per vertex         -> 32 V
    """
    n, v, t = st["corners"], st["vertices"], st["triangles"]
    words = n / 64
    algo = {
        "ms_plan": st["bricks_total"] * 24,
        "ms_corners": n / 8,
        "ms_classify": 4 * n / 8 + words * 8 * 3,
        "ms_emit": 4 * n / 8 + words * 8 + 8 * v + 12 * t,
        "ms_bisect": 32 * v,
    }[stage]
    gbs = algo / (ms * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": {"ms_plan": "k_plan", "ms_corners": "k_corners", "ms_classify": "k_classify",
                   "ms_emit": "k_emit", "ms_bisect": "k_bisect"}[stage],
        "achieved": gbs,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": gbs / HBM_PEAK_GBS,
        "traffic": None,
        "avg_ms": ms,
    }


if __name__ == "__main__":
    main()
