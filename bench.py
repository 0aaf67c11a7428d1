"""Benchmark: MarchingCubesSearch corner evals/s + tris/s on one MI355X.

A step is one call of the library on the workload, results left in HBM
(device_only); the solids are lowered and uploaded inside the step.

  csg64     (default) the seeded 64-primitive CSG tree clipped to [-1, 1]^3 at
            delta 0.005 (402 x 395 x 402 ~ 64M corners), iters 8: one
            MarchingCubesSearch (model3d/mc.go:45-49) per step
  sphere    Sphere{0, 1} at delta 0.02, iters 8
  collider  the CSG tree meshed at delta 0.02, as a ColliderSolid, 256 samples
            along its longest axis, iters 8
  batch     256 seeded CSG solids of 4..128 primitives on 32..256-cell grids,
            iters 4..8, all in one mc3d_marching_cubes_search_batch call
  large     a seeded 512-primitive CSG tree in [-1, 1]^3 at delta 0.002
            (1002^3 ~ 1.0G corners), iters 8

Prints one JSON line: the metric, the per-kernel times (HIP events on the
library's stream), the roofline of the dominant kernel (algorithmic flops or
bytes per launch over its average duration; HBM traffic per launch from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same
configuration), the guard statistics, the plain one-solid-at-a-time float64
GPU path, the end-to-end time with the copy to host, and the CPU baseline
(the oracle, a C restatement of the reference algorithm, on this host's
cores).
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
COPY_CALIBRATION_BYTES = 1 << 30
CPU_SAMPLE_SECONDS = 10.0

# The configuration each workload runs in, as chosen by model3d_amd.autotune
# on an MI355X (`--autotune` searches again): precision, culling, brick edge,
# grid rows per program walk, wavefronts per brick.
DEFAULT_RUN = {
    "csg64": dict(precision="fp64", cull=True, brick=8, rows_per_walk=1, waves_per_brick=4),
    "sphere": dict(precision="fp64", cull=False, brick=0, rows_per_walk=1, waves_per_brick=4),
    "collider": dict(precision="fp64", cull=False, brick=0, rows_per_walk=1, waves_per_brick=4),
    "batch": dict(precision="fp64", cull=True, brick=8, rows_per_walk=1, waves_per_brick=4),
    "large": dict(precision="fp64", cull=True, brick=8, rows_per_walk=1, waves_per_brick=4),
}
PLAIN_RUN = dict(precision="fp64", cull=False, brick=0, rows_per_walk=1, waves_per_brick=4)

# Kernel of each timed stage (mc3d_stats.ms_*), for the roofline and the
# rocprofv3 summaries.
STAGE_KERNEL = {
    "ms_plan": "k_plan",
    "ms_corners": "k_corners_act",
    "ms_classify": "k_classify",
    "ms_scan": "k_scan_*",
    "ms_emit": "k_emit",
    "ms_bisect": "k_bisect",
}
SUMMED = ("corners", "vertices", "triangles", "bisection_evals", "bricks_total", "program_length")


def workload(name):
    """{"solids", "deltas", "iters", "cfg"} of a named workload."""
    from model3d_amd import workloads

    if name == "csg64":
        c = workloads.CSG64_CONFIG
        return dict(solids=[workloads.csg64_solid(c["seed"], c["n_prims"])], deltas=[c["delta"]],
                    iters=[c["iters"]], cfg=c)
    if name == "sphere":
        c = workloads.SPHERE_CONFIG
        return dict(solids=[workloads.sphere_solid()], deltas=[c["delta"]], iters=[c["iters"]], cfg=c)
    if name == "collider":
        c = dict(workloads.COLLIDER_CONFIG)
        solid, delta, ntri = workloads.collider_solid(c["seed"], c["n_prims"], c["mesh_delta"])
        c["delta"], c["mesh_triangles"] = delta, ntri
        return dict(solids=[solid], deltas=[delta], iters=[c["iters"]], cfg=c)
    if name == "large":
        c = dict(workloads.LARGE_CONFIG, cpu_delta=0.008)
        return dict(solids=[workloads.large_solid(c["seed"], c["n_prims"])], deltas=[c["delta"]],
                    iters=[c["iters"]], cfg=c)
    if name == "batch":
        c = workloads.BATCH_CONFIG
        solids, deltas, iters = workloads.batch_solids(c["seed"], c["n_solids"], c["prims"], c["cells"], c["iters"])
        return dict(solids=solids, deltas=deltas, iters=iters, cfg=c)
    raise SystemExit(f"unknown workload {name}")


def precision_code(name):
    from model3d_amd import _lib

    return _lib.MC3D_PREC_FP32_GUARD if name == "fp32_guard" else _lib.MC3D_PREC_FP64


def merge_stats(per_mesh):
    """Call-level stats (timings, guard and operation counts are per call)
    with the per-solid counts summed and the occupancy words of all grids."""
    st = dict(per_mesh[0])
    for k in SUMMED:
        st[k] = sum(m[k] for m in per_mesh)
    st["words"] = sum((m["grid_nx"] + 63) // 64 * m["grid_ny"] * m["grid_nz"] for m in per_mesh)
    return st


def run_gpu(work, run, steps, warmup, count_mismatch=False, device_only=True, each=False, no_graph=False,
            copy_out=False):
    """`steps` timed steps; a step is one library call on all the solids (the
    batch entry point for several), or with `each` one call per solid.
    Results stay on the device unless device_only is False (then they arrive
    in the library's pinned memory, and with copy_out are also copied into
    caller arrays allocated once).  no_graph enqueues the stages one by one,
    which times each stage.  Returns (wall seconds of the timed steps, merged
    stats of each recorded step after them)."""
    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    L = _lib.lib()
    ctx = _lib.context(0)
    b = LibBackend()
    solids = work["solids"]
    n = len(solids)
    hs = [s._build(b) for s in solids]
    opts = _lib.mc3d_options()
    opts.precision = precision_code(run["precision"])
    opts.cull = 1 if run["cull"] else 0
    opts.brick, opts.rows_per_walk, opts.waves_per_brick = run["brick"], run["rows_per_walk"], run["waves_per_brick"]
    opts.device_only = 1 if device_only else 0
    opts.count_mismatch = 1 if count_mismatch else 0
    opts.no_graph = 1 if no_graph else 0
    opts.slab_planes = run.get("slab_planes", 0)
    bufs = {}

    def fetch(m):
        """The mesh into caller arrays (allocated on first use, then reused)."""
        import numpy as np

        nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
        L.mc3d_mesh_counts(m, ctypes.byref(nv), ctypes.byref(nt))
        if bufs.get("n", (-1, -1)) < (nv.value, nt.value):
            bufs["n"] = (nv.value, nt.value)
            bufs["v"] = np.empty((nv.value, 3))
            bufs["f"] = np.empty((nt.value, 3), dtype=np.uint32)
        _lib.check(L.mc3d_mesh_copy(m, bufs["v"].ctypes.data, bufs["f"].ctypes.data), "mc3d_mesh_copy")
    harr = (ctypes.c_void_p * n)(*hs)
    darr = (ctypes.c_double * n)(*work["deltas"])
    iarr = (ctypes.c_int32 * n)(*work["iters"])
    stats = []

    def mesh_stats(m):
        st = _lib.mc3d_stats()
        L.mc3d_mesh_stats(m, ctypes.byref(st))
        return st.as_dict()

    def step(record):
        per = []
        if n == 1 or each:
            # one call per solid, each mesh consumed (freed) before the next call
            for i in range(n):
                one = ctypes.c_void_p()
                _lib.check(L.mc3d_marching_cubes_search(ctx, hs[i], darr[i], iarr[i], ctypes.byref(opts),
                                                        ctypes.byref(one)), "mc3d_marching_cubes_search")
                if copy_out:
                    fetch(one)
                if record:
                    per.append(mesh_stats(one))
                L.mc3d_mesh_free(one)
        else:
            outs = (ctypes.c_void_p * n)()
            _lib.check(L.mc3d_marching_cubes_search_batch(ctx, harr, darr, iarr, n, ctypes.byref(opts), outs),
                       "mc3d_marching_cubes_search_batch")
            if record:
                per = [mesh_stats(outs[i]) for i in range(n)]
            for i in range(n):
                L.mc3d_mesh_free(outs[i])
        if record:
            merged = merge_stats(per)
            if n > 1 and each:  # one call per solid: its stage times add up
                for k in list(STAGE_KERNEL) + ["ms_device"]:
                    merged[k] = sum(m[k] for m in per)
            stats.append(merged)

    for _ in range(warmup):
        step(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(False)
    wall = time.perf_counter() - t0
    # The statistics come from further steps after the timed region: reading
    # them costs host time per mesh, which is not the call's.
    for _ in range(max(1, min(steps, 3))):
        step(True)
    for h in hs:
        b.release(h)
    b.close()
    return wall, stats


def mean(stats, key):
    return sum(s[key] for s in stats) / max(1, len(stats))


def host_threads():
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    return max(1, min(threads, 16))


def cpu_baseline(work):
    from oracle.binding import oracle_contains, oracle_search

    cfg = work["cfg"]
    threads = host_threads()
    if cfg.get("mesh_triangles"):
        # Brute-force ray parity over the whole mesh is far too slow for the
        # full grid on the CPU: time Contains on a sample of grid corners.
        import numpy as np

        solid = work["solids"][0]
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
    corners = ntris = 0
    wall = 0.0
    done = 0
    coarse = cfg.get("cpu_delta")  # the full grid takes the C restatement minutes: a coarser grid of the same solid
    for s, d, it in zip(work["solids"], work["deltas"], work["iters"]):
        d = coarse or d
        t0 = time.perf_counter()
        tris, info = oracle_search(s, d, it, threads=threads)
        wall += time.perf_counter() - t0
        corners += info["corners"]
        ntris += len(tris)
        done += 1
        if wall > CPU_SAMPLE_SECONDS:
            break
    n = len(work["solids"])
    what = "the full workload once" if done == n else f"the first {done} of the {n} solids"
    if coarse:
        what = f"the same solid at delta {coarse} instead of {work['deltas'][0]}"
    return {
        "value": corners / wall,
        "unit": "corner evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what} ({corners} corners, {ntris} triangles), C restatement of mc.go with its "
        f"plane-pipeline/parallel-search structure, {wall:.2f} s",
        "tris_per_s": ntris / wall,
    }


def algorithmic(stage, st):
    """Algorithmic work of one launch of a stage's kernel(s): ("flop", f64, f32)
    for the evaluation stages, ("bytes", B) for the grid / compaction stages.

    classify: reads the occupancy grid once (N/8 B) and writes a 4 B packed count
              per word
    scan    : reads the counts (4 B), writes the offsets (8 B) per word (the
              block sums and the 4 B entries of listed words are small)
    emit    : reads each listed word (8 B) and its 8 B offsets, and the 4 B list
              entry, writes an 8 B key per vertex and 12 B per triangle (the
              neighbour words it also reads are listed words' own or cached)
    plan    : writes an 8 B brick record per brick, a 4 B list entry per
              active brick and the 8 B occupancy words of the constant bricks
              (their share of the words taken as their share of the bricks)
    corners / bisect: the counted operations of every evaluated primitive,
              transform and midpoint (float64, float32)
    """
    n, v, t = st["corners"], st["vertices"], st["triangles"]
    words = st["words"]
    if stage == "ms_corners":
        return ("flop", st["corner_flops64"], st["corner_flops32"])
    if stage == "ms_bisect":
        return ("flop", st["bisect_flops64"], st["bisect_flops32"])
    b = {
        "ms_plan": 8 * st["bricks_total"] + 4 * st["bricks_active"]
        + 8 * words * (1 - st["bricks_active"] / max(1, st["bricks_total"])),
        "ms_classify": n / 8 + 4 * words,
        "ms_scan": 12 * words,
        "ms_emit": 20 * st["words_listed"] + 8 * v + 12 * t,
    }[stage]
    return ("bytes", b)


def roofline_of(stage, ms, st):
    kind = algorithmic(stage, st)
    if kind[0] == "flop":
        _, f64, f32 = kind
        # float64-equivalent rate: float32 runs at twice the float64 vector rate
        achieved = (f64 + f32 / 2) / (ms * 1e-3) / 1e12
        r = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
             "frac": achieved / FP64_PEAK_TFLOPS, "flops64_per_launch": f64, "flops32_per_launch": f32}
    else:
        achieved = kind[1] / (ms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": kind[1]}
    r["traffic"] = None
    r["kernel"] = STAGE_KERNEL[stage]
    r["avg_ms"] = ms
    return r


def load_traffic(workload_name, run, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    of this workload and configuration, or (None, reason)."""
    from model3d_amd.profiling import traffic_bytes

    path = os.path.join(ROOT, "profiles", f"{workload_name}_pmc_traffic.json")
    if not os.path.exists(path):
        return None, "no rocprofv3 FETCH_SIZE/WRITE_SIZE summary for this workload"
    s = json.load(open(path))
    if s.get("config") != run:
        return None, f"summary measured at {s.get('config')}, not at this configuration"
    names = [k for k in s.get("kernels", {}) if k.split("<")[0] == kernel]
    if len(names) > 1:  # instantiations of several precisions: the run's own (template argument MODE)
        mode = "1" if run["precision"] == "fp32_guard" else "0"
        names = [k for k in names if k.split("<")[1].split(",")[0].strip() == mode]
    if len(names) != 1:
        return None, f"kernel {kernel} not (uniquely) in the summary"
    return traffic_bytes(s, names[0]), os.path.relpath(path, ROOT)


def copy_calibration():
    """Two plain streams of COPY_CALIBRATION_BYTES for the FETCH_SIZE /
    WRITE_SIZE calibration of a rocprofv3 --pmc run: the library's
    k_stream_read reads them once, k_stream_write writes them once (each
    after an untimed filling pass).  Returns their bandwidths in GB/s."""
    from model3d_amd import _lib

    L = _lib.lib()
    ctx = _lib.context(0)
    out = {}
    for name, write in (("read", 0), ("write", 1)):
        ms = ctypes.c_double()
        _lib.check(L.mc3d_diag_stream(ctx, COPY_CALIBRATION_BYTES, write, ctypes.byref(ms)), "mc3d_diag_stream")
        out[name] = COPY_CALIBRATION_BYTES / (ms.value * 1e-3) / 1e9
    return out


def next_rows(work, run, steps=5):
    """The entry points next to MarchingCubesSearch on one solid of the
    workload, results left on the device, ms per call (after one warm-up):
    MarchingCubesInterior, MarchingCubesConj (a rotation and a translation),
    MarchingCubesSearchFilter with the native rect collider of the solid's
    mesh at 4x delta, MarchingCubesC2F (4x delta to delta) and the STL
    encoding of the mesh (from its device buffers)."""
    import numpy as np

    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend, Rotation, Translate

    L = _lib.lib()
    ctx = _lib.context(0)
    b = LibBackend()
    solid, delta, iters = work["solids"][0], work["deltas"][0], work["iters"][0]
    h = solid._build(b)
    o = _lib.mc3d_options()
    o.precision, o.cull, o.device_only = precision_code(run["precision"]), 1 if run["cull"] else 0, 1
    o.brick, o.rows_per_walk, o.waves_per_brick = run["brick"], run["rows_per_walk"], run["waves_per_brick"]
    out = {}

    def timed(name, call, extra=None):
        ms, st = [], None
        for i in range(steps + 1):
            m = ctypes.c_void_p()
            t0 = time.perf_counter()
            _lib.check(call(m), name)
            dt = time.perf_counter() - t0
            if i:
                ms.append(1e3 * dt)
            if st is None:
                s_ = _lib.mc3d_stats()
                L.mc3d_mesh_stats(m, ctypes.byref(s_))
                st = s_.as_dict()
            if extra and i == steps:
                extra(m, st)
            L.mc3d_mesh_free(m)
        r = {"ms_per_call": sum(ms) / len(ms), "corners": st["corners"], "triangles": st["triangles"],
             "corner_evals_per_s": st["corners"] / (sum(ms) / len(ms) / 1e3), "ms_device": st["ms_device"]}
        out[name] = r
        return r

    try:
        timed("interior", lambda m: L.mc3d_marching_cubes_interior(ctx, h, delta, iters, ctypes.byref(o), ctypes.byref(m)))
        xf = [op for t in (Rotation((0.0, 0.6, 0.8), 0.5), Translate((0.1, -0.2, 0.05))) for op in t._ops()]
        arr = (_lib.mc3d_xform * len(xf))()
        for i, (k, v) in enumerate(xf):
            arr[i].kind = k
            for j in range(9):
                arr[i].v[j] = v[j]
        timed("conj", lambda m: L.mc3d_marching_cubes_conj(ctx, h, delta, iters, arr, len(xf), ctypes.byref(o),
                                                            ctypes.byref(m)))
        # the solid's own mesh at 4x delta as the filter's collider
        coarse = ctypes.c_void_p()
        _lib.check(L.mc3d_marching_cubes_search(ctx, h, 4 * delta, iters, ctypes.byref(o), ctypes.byref(coarse)), "coarse")
        nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
        L.mc3d_mesh_counts(coarse, ctypes.byref(nv), ctypes.byref(nt))
        v = np.empty((nv.value, 3))
        f = np.empty((nt.value, 3), dtype=np.uint32)
        L.mc3d_mesh_copy(coarse, v.ctypes.data, f.ctypes.data)
        L.mc3d_mesh_free(coarse)
        tris = np.ascontiguousarray(v[f.astype(np.int64)].reshape(-1, 9))
        col = L.mc3d_rect_collider_create(tris.ctypes.data_as(_lib._dp), len(tris))
        fn = ctypes.cast(L.mc3d_rect_collider_collides, ctypes.c_void_p)
        r = timed("search_filter", lambda m: L.mc3d_marching_cubes_search_filter(ctx, h, fn, col, delta, iters,
                                                                                 ctypes.byref(o), ctypes.byref(m)))
        r["collider_triangles"] = len(tris)
        L.mc3d_rect_collider_free(col)
        timed("c2f", lambda m: L.mc3d_marching_cubes_c2f(ctx, h, 4 * delta, delta, 0.0, iters, ctypes.byref(o),
                                                         ctypes.byref(m)))

        def stl_and_group(m, st):
            n = L.mc3d_stl_bytes(st["triangles"])
            buf = np.empty(n, dtype=np.uint8)
            t = []
            for _ in range(steps):
                t0 = time.perf_counter()
                _lib.check(L.mc3d_mesh_encode_stl(ctx, m, buf.ctypes.data, n), "stl")
                t.append(1e3 * (time.perf_counter() - t0))
            out["stl"] = {"ms_per_encode": sum(t) / len(t), "bytes": n,
                          "GB_per_s_to_host": n / (sum(t) / len(t) / 1e3) / 1e9}
            # GroupTriangles (SaveGroupedSTL's order) of the mesh in place on
            # the device
            t = []
            for _ in range(steps):
                t0 = time.perf_counter()
                _lib.check(L.mc3d_mesh_group_triangles(ctx, m), "group")
                t.append(1e3 * (time.perf_counter() - t0))
            out["group_triangles"] = {"ms_per_call": sum(t) / len(t), "triangles": st["triangles"]}

        timed("search", lambda m: L.mc3d_marching_cubes_search(ctx, h, delta, iters, ctypes.byref(o), ctypes.byref(m)),
              extra=stl_and_group)
    finally:
        b.release(h)
        b.close()
    return out


def stage_times(work, run, steps=4):
    """Per-stage times (HIP events between the stages, graphs off) and the
    algorithmic counts of a statistics run, merged: (stages_ms, counts)."""
    _, st = run_gpu(work, run, steps, 1, no_graph=True)
    stages = {k: mean(st, k) for k in STAGE_KERNEL}
    stages["ms_device"] = mean(st, "ms_device")
    _, sst = run_gpu(work, run, 1, 0, count_mismatch=True)
    return stages, dict(sst[0])


def kernel_rooflines(stages, counts, run):
    kernels = {}
    for stage in STAGE_KERNEL:
        if stages[stage] > 0:
            kernels[STAGE_KERNEL[stage]] = roofline_of(stage, stages[stage], counts)
    if not run["cull"]:
        kernels.pop("k_corners_act", None)
        kernels["k_corners"] = dict(roofline_of("ms_corners", stages["ms_corners"], counts), kernel="k_corners")
    return kernels


def config_entry(name, steps=5, warmup=2):
    """A compact line for one of the other configurations: ms per step in the
    workload's configuration (results on the device, graph launches), in the
    other precision, corner evals/s and tris/s, and the dominant kernel's
    roofline from a stage-timed run."""
    work = workload(name)
    run = dict(DEFAULT_RUN[name])
    wall, st = run_gpu(work, run, steps, warmup)
    ms = 1e3 * wall / steps
    other = "fp32_guard" if run["precision"] == "fp64" else "fp64"
    wall2, _ = run_gpu(work, dict(run, precision=other), steps, 1)
    stages, counts = stage_times(work, run, steps=3)
    kernels = kernel_rooflines(stages, counts, run)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    s0 = st[0]
    e = {
        "workload": work["cfg"]["name"],
        "solids": len(work["solids"]),
        "corners": s0["corners"],
        "triangles": s0["triangles"],
        "ms_per_step": ms,
        f"ms_per_step_{other}": 1e3 * wall2 / steps,
        "corner_evals_per_s": s0["corners"] / (ms / 1e3),
        "tris_per_s": s0["triangles"] / (ms / 1e3),
        "ms_device": mean(st, "ms_device"),
        "dominant": {x: kernels[dom][x] for x in ("kernel", "avg_ms", "bound", "achieved", "unit", "frac")},
        **run,
    }
    if name == "large":
        # the same search streamed in z slabs of 125 cube layers (the grid's
        # device memory bounded by one slab; the meshes stitched)
        planes = 125
        wall3, st3 = run_gpu(work, dict(run, slab_planes=planes), max(2, steps // 2), 1)
        assert st3[0]["triangles"] == s0["triangles"] and st3[0]["vertices"] == s0["vertices"]
        e["z_slabs"] = {"slab_planes": planes, "slabs": -(-(s0["grid_nz"] - 1) // planes),
                        "ms_per_step": 1e3 * wall3 / max(2, steps // 2)}
    return e


def sphere_batched(copies=64, steps=5):
    """The north-star comparison on config 1: `copies` spheres (C1) meshed by
    one batch call in float32 + guard, against the plain float64 GPU path
    (every atom at every corner, one call per sphere), corner evals/s."""
    c = dict(workload("sphere"))
    work = dict(solids=c["solids"] * copies, deltas=c["deltas"] * copies, iters=c["iters"] * copies, cfg=c["cfg"])
    run = dict(DEFAULT_RUN["sphere"], precision="fp32_guard", cull=True, brick=8)
    wall, st = run_gpu(work, run, steps, 1)
    corners = st[0]["corners"]  # summed over the copies
    wall64, _ = run_gpu(work, PLAIN_RUN, max(1, steps // 2), 1, each=True)
    batched = corners / (wall / steps)
    plain = corners / (wall64 / max(1, steps // 2))
    return {"copies": copies, "corners_per_step": corners, "ms_per_step_batched_fp32_guard": 1e3 * wall / steps,
            "ms_per_step_plain_fp64": 1e3 * wall64 / max(1, steps // 2), "corner_evals_per_s_batched": batched,
            "corner_evals_per_s_plain_fp64": plain, "ratio": batched / plain}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="csg64", choices=sorted(DEFAULT_RUN))
    ap.add_argument("--autotune", action="store_true", help="search tile shapes and precision first")
    ap.add_argument("--profile-run", action="store_true",
                    help="only the warm-up and timed steps (for rocprofv3; stages enqueued one by one), plus the "
                         "calibration streams")
    ap.add_argument("--next-rows-only", action="store_true",
                    help="only the entry points next to the search (for rocprofv3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--no-configs", action="store_true", help="skip the lines of the other configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU only: the library runs on one GPU and has no multi-GPU path "
            "(sharding solids or z slabs over GPUs is a design only)",
            file=sys.stderr,
        )
        raise SystemExit(2)

    from model3d_amd import _lib  # noqa: F401  (fails loudly without the HIP library)

    work = workload(args.workload)
    cfg = work["cfg"]
    run = dict(DEFAULT_RUN[args.workload])
    tuned = None
    if args.autotune:
        from model3d_amd import autotune

        tuned = {}
        single = len(work["solids"]) == 1
        target = work["solids"][0] if single else work["solids"]
        deltas = work["deltas"][0] if single else work["deltas"]
        iters = work["iters"][0] if single else work["iters"]
        for label in ("fp32_guard", "fp64"):
            tuned[label] = autotune.tune(target, deltas, iters, precision=precision_code(label))
        best = min(tuned, key=lambda k: tuned[k]["ms"])
        run = dict(precision=best, cull=True, brick=tuned[best]["brick"], rows_per_walk=tuned[best]["rows_per_walk"],
                   waves_per_brick=tuned[best]["waves_per_brick"])

    if args.next_rows_only:
        print(json.dumps({"workload": args.workload, "config": run, "next_rows": next_rows(work, run, args.steps)}),
              flush=True)
        return

    if args.profile_run:
        streams = copy_calibration()
        wall, st = run_gpu(work, run, args.steps, args.warmup, no_graph=True)
        print(json.dumps({"workload": args.workload, "config": run, "ms_per_step": 1e3 * wall / args.steps,
                          "stages_ms": {k: mean(st, k) for k in STAGE_KERNEL}, "stream_GBps": streams}), flush=True)
        return

    # The timed steps: one call per step, results left in HBM, the chain of
    # stages launched as one captured HIP graph.
    wall, st = run_gpu(work, run, args.steps, args.warmup)
    s0 = st[0]
    corners, tris, verts = s0["corners"], s0["triangles"], s0["vertices"]
    ms = 1e3 * wall / args.steps
    nsolids = len(work["solids"])
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
        "dtype": "f32+f64" if run["precision"] == "fp32_guard" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{cfg['name']}_iters{cfg['iters']}" if nsolids == 1 else cfg["name"],
            "solids": nsolids,
            "corners": corners,
            "vertices": verts,
            "triangles": tris,
            "bisection_evals": s0["bisection_evals"],
            **run,
        },
        "tris_per_s": tris / (wall / args.steps),
        "ms_device": mean(st, "ms_device"),
    }
    # The same steps in both precisions (bit-identical results).
    other = "fp32_guard" if run["precision"] == "fp64" else "fp64"
    wall_o, _ = run_gpu(work, dict(run, precision=other), args.steps, args.warmup)
    result["precisions_ms_per_step"] = {run["precision"]: ms, other: 1e3 * wall_o / args.steps}
    if tuned:
        result["autotune"] = {k: {"brick": v["brick"], "rows_per_walk": v["rows_per_walk"],
                                  "waves_per_brick": v["waves_per_brick"], "best_ms": v["ms"]} for k, v in tuned.items()}

    # Stage times from a run with the stages enqueued one by one (HIP events
    # between them), and the algorithmic counts of a statistics run.
    stages, counts = stage_times(work, run)
    result["stages_ms"] = stages
    result["bricks"] = {"total": s0["bricks_total"], "active": s0["bricks_active"], "full": s0["bricks_full"],
                        "code_mean": s0["brick_code_total"] / max(1, s0["bricks_active"]),
                        "code_max": s0["brick_code_max"], "program": s0["program_length"]}

    # Guard statistics (float32 decisions checked against float64).
    if run["precision"] == "fp32_guard":
        g = counts
    else:
        _, gst = run_gpu(work, dict(run, precision="fp32_guard"), 1, 0, count_mismatch=True)
        g = gst[0]
    result["guard"] = {
        "fp32_atom_evals": g["guard_evals"],
        "fp64_rechecks": g["guard_rechecks"],
        "fp64_fallback_fraction": g["guard_rechecks"] / g["guard_evals"] if g["guard_evals"] else None,
        "fp32_sign_mismatch_rate": g["guard_mismatches"] / g["guard_evals"] if g["guard_evals"] else None,
        "band_violations": g["guard_violations"],
    }

    # Roofline of every kernel; the dominant one is the headline.
    kernels = kernel_rooflines(stages, counts, run)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    traffic, source = load_traffic(args.workload, run, dom.rstrip("*").rstrip("_"))
    head = dict(kernels[dom])
    head["traffic"] = traffic
    head["traffic_source"] = source
    result["roofline"] = head
    result["kernels_roofline"] = {k: {x: r[x] for x in ("bound", "achieved", "unit", "frac", "avg_ms")}
                                  for k, r in kernels.items()}

    # End to end through the C-ABI with the meshes fetched to the host: in
    # the library's pinned memory (read in place, as the Go drop-in does), and
    # left on the device, then copied by mc3d_mesh_copy into caller arrays
    # allocated once.
    n_e2e = max(1, args.steps // 2)
    wall_e2e, est = run_gpu(work, run, n_e2e, 1, device_only=False)
    wall_cp, _ = run_gpu(work, run, n_e2e, 1, device_only=True, copy_out=True)
    result["e2e_with_d2h"] = {"ms_per_step": 1e3 * wall_e2e / n_e2e, "ms_d2h": mean(est, "ms_d2h"),
                              "ms_per_step_into_caller_arrays": 1e3 * wall_cp / n_e2e}

    if not args.no_fp64:
        # the plain float64 GPU path: every atom at every corner, one call per solid
        n64 = max(1, min(args.steps // 2, 3))
        wall64, st64 = run_gpu(work, PLAIN_RUN, n64, 1, each=True)
        ms64 = 1e3 * wall64 / n64
        result["fp64_plain"] = {
            "corner_evals_per_s": corners / (ms64 / 1e3),
            "tris_per_s": tris / (ms64 / 1e3),
            "ms_per_step": ms64,
        }
        result["speedup_vs_fp64_plain"] = ms64 / ms
        if nsolids > 1:
            # the same configuration, one call per solid
            walle, _ = run_gpu(work, run, n64, 1, each=True)
            result["one_call_per_solid"] = {"ms_per_step": 1e3 * walle / n64,
                                            "speedup_of_batch": (1e3 * walle / n64) / ms}

    if nsolids == 1 and args.workload in ("csg64", "sphere"):
        result["next_rows"] = next_rows(work, run)

    if not args.no_configs:
        result["configs"] = {name: config_entry(name) for name in ("sphere", "collider", "batch", "large")
                             if name != args.workload}
        result["sphere_batched"] = sphere_batched()

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(work)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
