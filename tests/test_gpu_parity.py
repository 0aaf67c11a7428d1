"""GPU parity: the HIP library against the oracle, through the C-ABI.

Bit-exact comparisons throughout: Contains booleans, and triangle sets as
multisets of 9 float64 bit patterns per triangle.
"""
import math

import numpy as np
import pytest

import model3d_amd as m3
from model3d_amd import workloads
from oracle.binding import oracle_contains, oracle_search
from tests.helpers import canonical, edge_topology, same_triangles, solid_catalogue, triangle_checksum, volume

pytestmark = pytest.mark.gpu

FP64 = m3.MC3D_PREC_FP64
GUARD = m3.MC3D_PREC_FP32_GUARD


def _near_surface_points(s, rng, n=4000):
    """Random points plus points on the bisection paths of the oracle mesh,
    which straddle the surface at all scales."""
    mn, mx = np.array(s.Min()), np.array(s.Max())
    pad = 0.1 * (mx - mn) + 1e-3
    p = rng.uniform(mn - pad, mx + pad, size=(n, 3))
    tris, _ = oracle_search(s, 0.05, 12, threads=8)
    v = tris.reshape(-1, 3)[: 3 * n]
    return np.concatenate([p, v, v + rng.normal(scale=1e-9, size=v.shape)])


@pytest.mark.parametrize("name", list(solid_catalogue()))
@pytest.mark.parametrize("precision", [FP64, GUARD])
def test_contains_parity(name, precision):
    s = solid_catalogue()[name]
    rng = np.random.default_rng(abs(hash(name)) % 1000)
    pts = _near_surface_points(s, rng)
    print(f"contains parity {name} precision={precision} n={len(pts)}", flush=True)
    got = m3.Contains(s, pts, precision=precision)
    want = oracle_contains(s, pts)
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} mismatches"


@pytest.mark.parametrize("name", list(solid_catalogue()))
@pytest.mark.parametrize("precision", [FP64, GUARD])
def test_mesh_parity_catalogue(name, precision):
    s = solid_catalogue()[name]
    print(f"mesh parity {name} precision={precision}", flush=True)
    want, _ = oracle_search(s, 0.04, 8, threads=8)
    got = m3.MarchingCubesSearch(s, 0.04, 8, precision=precision)
    assert len(want) > 0
    assert same_triangles(got.triangles(), want)
    # one wave per brick, four rows per walk, small bricks
    got = m3.MarchingCubesSearch(s, 0.04, 8, precision=precision, brick=2, rows_per_walk=4, waves_per_brick=1)
    assert same_triangles(got.triangles(), want)


@pytest.mark.parametrize("iters", [0, 1, 8])
def test_sphere_config_bit_exact(iters):
    s = workloads.sphere_solid()
    print(f"sphere r=1 d=0.02 iters={iters}", flush=True)
    want, info = oracle_search(s, 0.02, iters, threads=8)
    for precision in (FP64, GUARD):
        got = m3.MarchingCubesSearch(s, 0.02, iters, precision=precision)
        assert got.stats["corners"] == 102**3
        assert got.stats["triangles"] == 93992 and got.stats["vertices"] == 46998
        assert same_triangles(got.triangles(), want)


@pytest.mark.parametrize("delta,tiles", [(0.013885, 64), (0.01379, 65)])
def test_scan_one_pass_boundary(delta, tiles):
    """Grids of 64 classify blocks (classify scans its counts itself) and 65
    (the separate scan), without culling: the meshes equal the oracle's."""
    s = m3.Sphere((0.0, 0.0, 0.0), 1.0)
    print(f"sphere d={delta} ({tiles} classify blocks)", flush=True)
    want, _ = oracle_search(s, delta, 4, threads=8)
    got = m3.MarchingCubesSearch(s, delta, 4, cull=False)
    st = got.stats
    words = (st["grid_nx"] + 63) // 64 * st["grid_ny"] * st["grid_nz"]
    assert ((words + 63) // 64 * 64 + 1023) // 1024 == tiles
    assert same_triangles(got.triangles(), want)


def test_csg64_coarse_bit_exact():
    s = workloads.csg64_solid(2)
    print("csg64 d=0.02", flush=True)
    want, _ = oracle_search(s, 0.02, 8, threads=8)
    for precision in (FP64, GUARD):
        got = m3.MarchingCubesSearch(s, 0.02, 8, precision=precision)
        assert same_triangles(got.triangles(), want)


def test_csg64_full_size_bit_exact():
    """The benchmark workload at its full size (delta 0.005, ~65M corners)."""
    s = workloads.csg64_solid(2)
    print("csg64 d=0.005 (full size)", flush=True)
    want, info = oracle_search(s, 0.005, 8)
    got = m3.MarchingCubesSearch(s, 0.005, 8, precision=GUARD)
    assert got.stats["corners"] == info["corners"]
    assert same_triangles(got.triangles(), want)
    got64 = m3.MarchingCubesSearch(s, 0.005, 8, precision=FP64)
    assert np.array_equal(canonical(got64.triangles()), canonical(got.triangles()))


def test_culling_and_brick_sizes_do_not_change_results():
    s = workloads.csg64_solid(2)
    print("csg64 d=0.01 cull/brick sweep", flush=True)
    ref = canonical(m3.MarchingCubesSearch(s, 0.01, 8, precision=FP64, cull=False).triangles())
    for precision in (FP64, GUARD):
        for cull, brick, rpw, wpb in ((False, 0, 0, 0), (True, 1, 1, 4), (True, 2, 2, 1), (True, 4, 4, 2),
                                      (True, 8, 1, 1), (True, 16, 4, 4), (True, 3, 2, 2), (True, 4, 2, 1),
                                      (True, 8, 4, 1), (True, 16, 1, 1), (True, 32, 2, 4)):
            got = m3.MarchingCubesSearch(s, 0.01, 8, precision=precision, cull=cull, brick=brick, rows_per_walk=rpw,
                                         waves_per_brick=wpb)
            assert np.array_equal(canonical(got.triangles()), ref), (precision, cull, brick, rpw, wpb)


@pytest.mark.parametrize("name", ["csg64", "sphere", "capsule", "cylinder", "csg", "transform_chain"])
def test_guard_band_never_wrong(name):
    """Every float32 decision the band calls sure agrees with float64."""
    s = workloads.csg64_solid(2) if name == "csg64" else solid_catalogue()[name]
    print(f"guard statistics {name}", flush=True)
    got = m3.MarchingCubesSearch(s, 0.01, 8, precision=GUARD, count_mismatch=True)
    st = got.stats
    assert st["guard_evals"] > 0
    assert st["guard_violations"] == 0
    assert st["guard_rechecks"] < st["guard_evals"]
    plain = m3.MarchingCubesSearch(s, 0.01, 8, precision=FP64)
    assert np.array_equal(canonical(got.triangles()), canonical(plain.triangles()))


def _search_on(ctx, solid, delta, iters):
    """MarchingCubesSearch through the C-ABI on an explicit context."""
    import ctypes

    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    L = _lib.lib()
    b = LibBackend()
    h = solid._build(b)
    out = ctypes.c_void_p()
    try:
        _lib.check(L.mc3d_marching_cubes_search(ctx, h, delta, iters, None, ctypes.byref(out)))
        nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
        L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt))
        v = np.empty((nv.value, 3)), np.empty((nt.value, 3), dtype=np.uint32)
        _lib.check(L.mc3d_mesh_copy(out, v[0].ctypes.data, v[1].ctypes.data))
        return v[0][v[1].astype(np.int64)]
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)


def test_solid_handle_reused_across_calls_and_threads():
    """A handle keeps its lowered program after the first search: later calls
    on it, from one thread or from four at once (each on its own context),
    give the oracle's mesh bit for bit."""
    import ctypes
    import threading

    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    s = workloads.csg64_solid(2)
    delta, iters = 0.03, 8
    print("csg64 d=0.03, one handle, serial then 4 threads", flush=True)
    want, _ = oracle_search(s, delta, iters, threads=8)
    L = _lib.lib()
    b = LibBackend()
    h = s._build(b)

    def search(ctx):
        out = ctypes.c_void_p()
        opts = _lib.mc3d_options()
        opts.precision = FP64
        opts.cull = 1
        try:
            _lib.check(L.mc3d_marching_cubes_search(ctx, h, delta, iters, ctypes.byref(opts), ctypes.byref(out)))
            nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
            L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt))
            v, t = np.empty((nv.value, 3)), np.empty((nt.value, 3), dtype=np.uint32)
            _lib.check(L.mc3d_mesh_copy(out, v.ctypes.data, t.ctypes.data))
            return v[t.astype(np.int64)]
        finally:
            if out.value:
                L.mc3d_mesh_free(out)

    ctxs = []
    try:
        for _ in range(3):
            assert same_triangles(search(_lib.context(0)), want)
        for _ in range(4):
            c = ctypes.c_void_p()
            _lib.check(L.mc3d_ctx_create(0, ctypes.byref(c)))
            ctxs.append(c)
        got, errs = [None] * 4, []

        def run(i):
            try:
                got[i] = search(ctxs[i])
            except Exception as e:  # reported below, on the test's thread
                errs.append(e)

        threads = [threading.Thread(target=run, args=(i,)) for i in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not errs, errs
        for g in got:
            assert g is not None and same_triangles(g, want)
    finally:
        for c in ctxs:
            L.mc3d_ctx_destroy(c)
        b.release(h)


def test_output_capacity_growth_and_reuse():
    """A fresh context starts with small output buffers: the first large mesh
    overflows them and the output stages run again; later calls reuse them."""
    import ctypes

    from model3d_amd import _lib

    L = _lib.lib()
    ctx = ctypes.c_void_p()
    _lib.check(L.mc3d_ctx_create(0, ctypes.byref(ctx)))
    try:
        big = workloads.csg64_solid(2)
        print("fresh context: csg64 d=0.005, sphere, csg64 again", flush=True)
        want_big = canonical(m3.MarchingCubesSearch(big, 0.005, 8).triangles())
        assert np.array_equal(canonical(_search_on(ctx, big, 0.005, 8)), want_big)
        small = m3.Sphere((0, 0, 0), 0.5)
        want_small = canonical(m3.MarchingCubesSearch(small, 0.05, 8).triangles())
        assert np.array_equal(canonical(_search_on(ctx, small, 0.05, 8)), want_small)
        assert np.array_equal(canonical(_search_on(ctx, big, 0.005, 8)), want_big)
    finally:
        L.mc3d_ctx_destroy(ctx)


def test_collider_of_csg_mesh_bit_exact():
    """A ColliderSolid over a 12k-triangle mesh of the CSG workload, meshed again."""
    s = workloads.csg64_solid(2)
    tris = m3.MarchingCubesSearch(s, 0.04, 8).triangles()
    col = m3.NewColliderSolid(m3.MeshCollider(tris))
    print(f"collider over {len(tris)} triangles at d=0.05", flush=True)
    want, _ = oracle_search(col, 0.05, 8)
    got = m3.MarchingCubesSearch(col, 0.05, 8)
    assert same_triangles(got.triangles(), want)
    # the meshed collider approximates the solid: volumes agree to the grid scale
    assert abs(volume(got.triangles()) - volume(tris)) < 0.05 * abs(volume(tris))


def test_mesh_topology_and_volume():
    s = m3.Sphere((0.1, 0.3, -0.2), 1.0)
    got = m3.MarchingCubesSearch(s, 0.05, 8)
    directed_once, paired = edge_topology(got.triangles())
    assert directed_once and paired
    assert abs(volume(got.triangles()) - 4 / 3 * math.pi) < 2e-2
    # every vertex is referenced, and indices are in range
    assert got.faces.max() < len(got.vertices)
    assert len(np.unique(got.faces)) == len(got.vertices)


class _LyingBounds(m3.Solid):
    """A host-evaluated solid that is true beyond the bounds it reports."""

    def Min(self):
        return (-0.3, -0.3, -0.3)

    def Max(self):
        return (0.3, 0.3, 0.3)

    def ContainsBatch(self, pts):
        return np.linalg.norm(np.asarray(pts), axis=1) <= 1.0


def test_inside_boundary_panics():
    forced = m3.ForceSolidBounds(m3.Sphere((0, 0, 0), 0.5), (-0.4, -0.4, -0.4), (0.4, 0.4, 0.4))
    ok = m3.MarchingCubesSearch(forced, 0.1, 2)
    assert ok.NumTriangles() > 0
    with pytest.raises(m3.Panic) as e:
        m3.MarchingCubesSearch(_LyingBounds(), 0.1, 0)
    assert "solid is true outside of bounds" in str(e.value)


class _Failing(m3.Solid):
    def Min(self):
        return (-0.5, -0.5, -0.5)

    def Max(self):
        return (0.5, 0.5, 0.5)

    def ContainsBatch(self, pts):
        raise RuntimeError("user solid failed")


def test_host_callback_error_is_reported():
    with pytest.raises(m3.MC3DError) as e:
        m3.MarchingCubesSearch(m3.JoinedSolid([m3.Sphere((0, 0, 0), 0.3), _Failing()]), 0.1, 2)
    assert "callback" in str(e.value)
    with pytest.raises(m3.MC3DError):
        m3.Contains(_Failing(), np.zeros((4, 3)))
    # the context is still usable afterwards
    assert m3.MarchingCubesSearch(m3.Sphere((0, 0, 0), 0.3), 0.1, 2).NumTriangles() > 0


def test_host_solid_batches_and_rounds():
    """A host-evaluated solid on a grid of more than one callback batch
    (2^20 points), with the bisection rounds answered by the host."""
    from tests.helpers import Superellipsoid

    s = m3.SubtractedSolid(Superellipsoid((0.0, 0.0, 0.0), 0.9), m3.Sphere((0.2, 0.1, 0.0), 0.5))
    print("host solid d=0.015 (>1M corners)", flush=True)
    meshes = []
    for precision in (FP64, GUARD):
        got = m3.MarchingCubesSearch(s, 0.015, 3, precision=precision)
        assert got.stats["corners"] > (1 << 20)
        assert got.stats["ms_host_callbacks"] > 0
        meshes.append(canonical(got.triangles()))
    assert np.array_equal(meshes[0], meshes[1])
    directed_once, paired = edge_topology(got.triangles())
    assert directed_once and paired
    # The oracle asks the callback one point at a time: compare on a coarser grid.
    want, _ = oracle_search(s, 0.05, 3, threads=8)
    assert same_triangles(m3.MarchingCubesSearch(s, 0.05, 3).triangles(), want)


def test_invalid_bounds_panics():
    with pytest.raises(m3.Panic) as e:
        m3.MarchingCubesSearch(m3.Sphere((0, 0, 0), float("nan")), 0.1, 0)
    assert "invalid bounds" in str(e.value)


def test_large_workload_properties():
    """The 1G-corner workload (512 primitives, delta 0.002): float32 guard and
    float64 agree bit for bit, every triangle index is in range, and the mesh
    is a closed, consistently oriented surface (each directed edge once, with
    its reverse) -- checked on the device's exact vertex ids."""
    c = workloads.LARGE_CONFIG
    s = workloads.large_solid(c["seed"], c["n_prims"])
    print("large: 512 primitives at delta 0.002", flush=True)
    a = m3.MarchingCubesSearch(s, c["delta"], c["iters"], precision=GUARD)
    assert a.stats["corners"] > 10**9 and a.stats["triangles"] > 10**7
    b = m3.MarchingCubesSearch(s, c["delta"], c["iters"], precision=FP64)
    assert np.array_equal(a.vertices.view(np.uint64), b.vertices.view(np.uint64))
    assert np.array_equal(a.faces, b.faces)
    f = a.faces.astype(np.int64)
    assert f.max() < len(a.vertices)
    nv = len(a.vertices)
    e = np.stack([f, np.roll(f, -1, axis=1)], axis=2).reshape(-1, 2)
    key = e[:, 0] * nv + e[:, 1]
    rev = e[:, 1] * nv + e[:, 0]
    key.sort()
    assert np.all(key[1:] != key[:-1])  # every directed edge once
    assert np.all(np.isin(rev, key))  # and its reverse is there
    # The same triangles through the other classify paths: z slabs (each
    # slab small enough for per-lane loads) and the unculled word order.
    want = triangle_checksum(b.triangles())
    print("large: z slabs of 125 layers, then unculled", flush=True)
    slabs = m3.MarchingCubesSearch(s, c["delta"], c["iters"], precision=FP64, slab_planes=125)
    assert len(slabs.faces) == len(b.faces) and triangle_checksum(slabs.triangles()) == want
    whole = m3.MarchingCubesSearch(s, c["delta"], c["iters"], precision=FP64, cull=False)
    assert len(whole.faces) == len(b.faces) and triangle_checksum(whole.triangles()) == want


def test_go_shim_call_sequence():
    """The calls the Go drop-in makes for MarchingCubesSearch, in order:
    lower the solid, search (FP64, culled, results fetched to pinned host
    memory), read the counts and the host arrays in place, free; and the
    same with no options (the library's defaults)."""
    import ctypes

    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    L = _lib.lib()
    ctx = _lib.context(0)
    s = workloads.csg64_solid(2)
    print("go shim sequence: csg64 d=0.01", flush=True)
    want, _ = oracle_search(s, 0.01, 8, threads=8)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    shim = _lib.mc3d_options()
    shim.precision, shim.cull, shim.device_only = FP64, 1, 0
    pv, pt = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        for opts in (ctypes.byref(shim), None):
            _lib.check(L.mc3d_marching_cubes_search(ctx, h, 0.01, 8, opts, ctypes.byref(out)), "search")
            nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
            _lib.check(L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt)), "counts")
            _lib.check(L.mc3d_mesh_host_ptrs(out, ctypes.byref(pv), ctypes.byref(pt)), "host ptrs")
            v = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_double)), shape=(nv.value, 3))
            f = np.ctypeslib.as_array(ctypes.cast(pt, ctypes.POINTER(ctypes.c_uint32)), shape=(nt.value, 3))
            assert same_triangles(v[f.astype(np.int64)], want)
            st = _lib.mc3d_stats()
            L.mc3d_mesh_stats(out, ctypes.byref(st))
            assert st.ms_d2h >= 0 and st.triangles == len(want)
            L.mc3d_mesh_free(out)
            out = ctypes.c_void_p()
        # a mesh left on the device has no host arrays
        dev = ctypes.c_void_p()
        o = _lib.mc3d_options()
        o.precision, o.cull, o.device_only = GUARD, 1, 1
        _lib.check(L.mc3d_marching_cubes_search(ctx, h, 0.05, 2, ctypes.byref(o), ctypes.byref(dev)), "search")
        assert L.mc3d_mesh_host_ptrs(dev, ctypes.byref(pv), ctypes.byref(pt)) == _lib.MC3D_E_INVALID_ARG
        L.mc3d_mesh_free(dev)
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)


@pytest.mark.gpu
def test_device_mesh_copy_into_caller_arrays():
    """mc3d_mesh_copy of a mesh left on the device (chunked through pinned
    buffers, more chunks than the ring holds) and of its interior points
    equals the same call's mesh fetched to the host, byte for byte, at the
    benchmark's size (19.5 MB of mesh)."""
    import ctypes

    from model3d_amd import _lib
    from model3d_amd.solids import LibBackend

    L = _lib.lib()
    ctx = _lib.context(0)
    s = workloads.csg64_solid(2)
    print("device mesh copy: csg64 d=0.005", flush=True)
    b = LibBackend()
    h = s._build(b)
    got = {}
    try:
        for dev in (0, 1):
            o = _lib.mc3d_options()
            o.precision, o.cull, o.device_only = FP64, 1, dev
            out = ctypes.c_void_p()
            _lib.check(L.mc3d_marching_cubes_interior(ctx, h, 0.005, 8, ctypes.byref(o), ctypes.byref(out)), "interior")
            try:
                nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
                _lib.check(L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt)), "counts")
                v = np.full((nv.value, 3), np.nan)
                f = np.full((nt.value, 3), 0xFFFFFFFF, dtype=np.uint32)
                inner = np.full((nv.value, 3), np.nan)
                _lib.check(L.mc3d_mesh_copy(out, v.ctypes.data, f.ctypes.data), "copy")
                _lib.check(L.mc3d_mesh_copy_interior(out, inner.ctypes.data), "copy interior")
                # either array alone
                f2 = np.zeros_like(f)
                _lib.check(L.mc3d_mesh_copy(out, None, f2.ctypes.data), "copy tris")
                assert np.array_equal(f, f2)
                got[dev] = (v, f, inner)
            finally:
                L.mc3d_mesh_free(out)
        assert got[0][0].nbytes + got[0][1].nbytes > 16 << 20
        for a, c in zip(got[0], got[1]):
            assert a.tobytes() == c.tobytes()
    finally:
        b.release(h)


def _spacer(lo, hi, delta):
    """newSquareSpacer's axis (mc.go:594-608): plain float adds, as Go's."""
    out, x = [], lo - delta
    while x <= hi + delta:
        out.append(x)
        x += delta
    return np.array(out)


def test_collider_workload_at_bench_size():
    """The collider workload the benchmark times: the CSG tree meshed at
    delta 0.02 (about 50k triangles) as a ColliderSolid, 256 samples along
    its longest axis, iters 8.  Contains is compared with the oracle's brute
    force over every triangle on 24k points -- grid corners on and off the
    surface, first bisection midpoints, searched vertices and points a hair
    away from them -- and the full grid is checked through properties: float32
    guard and float64 agree bit for bit, the surface is closed and
    consistently oriented, and its volume matches the source mesh's to the
    grid scale (ColliderContains, collisions.go:119-134)."""
    c = workloads.COLLIDER_CONFIG
    solid, delta, ntri = workloads.collider_solid(c["seed"], c["n_prims"], c["mesh_delta"])
    print(f"collider bench size: {ntri} triangles, delta {delta}", flush=True)
    assert ntri > 40000
    a = m3.MarchingCubesSearch(solid, delta, c["iters"], precision=GUARD)
    b = m3.MarchingCubesSearch(solid, delta, c["iters"], precision=FP64)
    assert max(a.stats["grid_nx"], a.stats["grid_ny"], a.stats["grid_nz"]) == 256
    assert a.stats["triangles"] > 50000
    assert np.array_equal(a.vertices.view(np.uint64), b.vertices.view(np.uint64))
    assert np.array_equal(a.faces, b.faces)

    rng = np.random.default_rng(3)
    mn, mx = solid.Min(), solid.Max()
    ax = [_spacer(mn[i], mx[i], delta) for i in range(3)]
    assert [len(x) for x in ax] == [a.stats["grid_nx"], a.stats["grid_ny"], a.stats["grid_nz"]]
    # grid corners: random ones, and the two ends of the edges the vertices lie on
    idx = np.stack([rng.integers(0, len(ax[i]), 6000) for i in range(3)], axis=1)
    corners = np.stack([ax[i][idx[:, i]] for i in range(3)], axis=1)
    v = b.vertices[rng.choice(len(b.vertices), 4000, replace=False)]
    ends, mids = [], []
    for p in v:
        k = [int(np.searchsorted(ax[i], p[i])) for i in range(3)]
        on = [i for i in range(3) if k[i] >= len(ax[i]) or ax[i][k[i]] != p[i]]
        assert len(on) == 1  # a vertex lies on one grid edge: two coordinates are grid values
        axis = on[0]
        lo, hi = ax[axis][k[axis] - 1], ax[axis][k[axis]]
        q0, q1, qm = p.copy(), p.copy(), p.copy()
        q0[axis], q1[axis], qm[axis] = lo, hi, (lo + hi) / 2  # the edge's corners and first midpoint
        ends += [q0, q1]
        mids.append(qm)
    near = v + rng.normal(scale=1e-9, size=v.shape)
    pts = np.concatenate([corners, np.array(ends), np.array(mids), v, near])
    assert len(pts) >= 20000
    want = oracle_contains(solid, pts, threads=16)
    assert 0 < want.sum() < len(pts)
    for precision in (FP64, GUARD):
        got = m3.Contains(solid, pts, precision=precision)
        assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} mismatches (precision {precision})"
    # every edge end pair straddles the surface, as the mesh says
    e = want[len(corners) : len(corners) + 2 * len(v)].reshape(-1, 2)
    assert np.all(e[:, 0] != e[:, 1])

    f = b.faces.astype(np.int64)
    nv = len(b.vertices)
    ed = np.stack([f, np.roll(f, -1, axis=1)], axis=2).reshape(-1, 2)
    key = ed[:, 0] * nv + ed[:, 1]
    rev = ed[:, 1] * nv + ed[:, 0]
    key.sort()
    assert np.all(key[1:] != key[:-1])  # every directed edge once
    assert np.all(np.isin(rev, key))  # and its reverse is there
    src, got_v = volume(solid.collider.triangles), volume(b.triangles())
    print(f"collider volume {got_v} against the source mesh's {src}", flush=True)
    assert abs(got_v - src) < 0.02 * abs(src), (got_v, src)


# ---------------------------------------------------------------------------
# z-slab streaming (mc.go:614-645 keeps a few planes alive; the library
# bounds the grid's device memory by one slab and stitches the meshes)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["csg64", "sphere", "torus", "collider_inset", "host_superellipsoid"])
@pytest.mark.parametrize("planes", [1, 5, 16])
def test_z_slabs_equal_whole_grid(name, planes):
    """The stitched mesh is the whole grid's: the same triangles bit for bit
    and the same number of vertices (each shared plane's vertices once)."""
    s = workloads.csg64_solid(2) if name == "csg64" else solid_catalogue()[name]
    delta = 0.02 if name == "csg64" else 0.04
    print(f"z slabs {name} planes={planes}", flush=True)
    whole = m3.MarchingCubesSearch(s, delta, 8, precision=FP64)
    got = m3.MarchingCubesSearch(s, delta, 8, precision=FP64, slab_planes=planes)
    assert got.stats["corners"] == whole.stats["corners"]
    assert len(got.vertices) == len(whole.vertices)
    assert np.array_equal(canonical(got.triangles()), canonical(whole.triangles()))
    assert len(np.unique(got.faces)) == len(got.vertices)  # every vertex used
    g = m3.MarchingCubesSearch(s, delta, 8, precision=GUARD, slab_planes=planes)
    assert np.array_equal(canonical(g.triangles()), canonical(whole.triangles()))


def test_z_slabs_interior_and_panic():
    s = solid_catalogue()["csg"]
    mesh, interior = m3.MarchingCubesInterior(s, 0.04, 5)
    smesh, sinterior = m3.MarchingCubesInterior(s, 0.04, 5, slab_planes=3)
    pairs = lambda v, p: {tuple(r) for r in np.concatenate([v, p], axis=1).view(np.uint64).tolist()}
    assert pairs(smesh.vertices, sinterior) == pairs(mesh.vertices, interior)
    # a solid true on the grid's top face still panics; its inner slab
    # planes are not faces
    for planes in (0, 4):
        with pytest.raises(m3.Panic, match="solid is true outside of bounds"):
            m3.MarchingCubesSearch(_TrueAbove(), 0.05, 2, slab_planes=planes)


class _TrueAbove(m3.Solid):
    """True for z > 0.2, also on the top face of its grid (z = 0.55) and
    on the inner planes of every slab above 0.2."""

    def Min(self):
        return (-0.5, -0.5, -0.5)

    def Max(self):
        return (0.5, 0.5, 0.5)

    def ContainsBatch(self, pts):
        return np.asarray(pts)[:, 2] > 0.2
