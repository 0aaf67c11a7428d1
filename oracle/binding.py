"""ctypes binding of oracle/liboracle.so (test infrastructure only).

``OracleBackend`` has the same constructor methods as the library backend in
model3d_amd.solids, so one solid tree built with the model3d_amd types can be
lowered to both and the outputs compared.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_D3 = ctypes.c_double * 3
_P = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
HOST_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp)
RECT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, _dp)

_SIG = {
    "or_sphere": (_P, [_D3, ctypes.c_double]),
    "or_rect": (_P, [_D3, _D3]),
    "or_capsule": (_P, [_D3, _D3, ctypes.c_double]),
    "or_cylinder": (_P, [_D3, _D3, ctypes.c_double]),
    "or_cone": (_P, [_D3, _D3, ctypes.c_double]),
    "or_cone_slice": (_P, [_D3, _D3, ctypes.c_double, ctypes.c_double]),
    "or_torus": (_P, [_D3, _D3, ctypes.c_double, ctypes.c_double]),
    "or_joined": (_P, [ctypes.POINTER(_P), ctypes.c_int]),
    "or_intersected": (_P, [ctypes.POINTER(_P), ctypes.c_int]),
    "or_subtracted": (_P, [_P, _P]),
    "or_transform": (_P, [_P, ctypes.POINTER(ctypes.c_int), _dp, ctypes.c_int]),
    "or_force_bounds": (_P, [_P, _D3, _D3]),
    "or_collider": (_P, [_dp, ctypes.c_long]),
    "or_collider_inset": (_P, [_dp, ctypes.c_long, ctypes.c_double]),
    "or_collider_hollow": (_P, [_dp, ctypes.c_long, ctypes.c_double]),
    "or_tri_sphere_collision": (ctypes.c_int, [_dp, _D3, ctypes.c_double]),
    "or_group_triangles": (ctypes.c_int, [_dp, ctypes.c_long, ctypes.c_void_p]),
    "or_host": (_P, [_D3, _D3, HOST_FN, ctypes.c_void_p]),
    "or_free": (None, [_P]),
    "or_bounds": (None, [_P, _D3, _D3]),
    "or_contains": (ctypes.c_int, [_P, _D3]),
    "or_contains_batch": (None, [_P, _dp, ctypes.c_long, ctypes.c_void_p]),
    "or_case_table": (ctypes.c_int, [ctypes.c_void_p]),
    "or_rotations": (ctypes.c_int, [ctypes.c_void_p]),
    "or_spacer": (ctypes.c_long, [_P, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "or_marching_cubes_search": (
        ctypes.c_int,
        [_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)],
    ),
    "or_marching_cubes_interior": (
        ctypes.c_int,
        [_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)],
    ),
    "or_mesh_interior": (_dp, [_P, ctypes.POINTER(ctypes.c_long)]),
    "or_inverse_apply": (None, [ctypes.POINTER(ctypes.c_int), _dp, ctypes.c_int, _dp, ctypes.c_long]),
    "or_encode_stl": (None, [_dp, ctypes.c_long, ctypes.c_void_p]),
    "or_marching_cubes_filter": (
        ctypes.c_int,
        [_P, RECT_FN, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)],
    ),
    "or_tris_rect_collision": (ctypes.c_int, [_dp, ctypes.c_long, _D3, _D3]),
    "or_mesh_ntris": (ctypes.c_long, [_P]),
    "or_mesh_tris": (_dp, [_P]),
    "or_mesh_info": (
        None,
        [_P] + [ctypes.POINTER(ctypes.c_long)] * 3 + [ctypes.POINTER(ctypes.c_double)] * 2,
    ),
    "or_mesh_free": (None, [_P]),
}

OR_E_INVALID_BOUNDS = 1
OR_E_INSIDE_BOUNDARY = 2
OR_E_NOT_ON_EDGE = 3
OR_E_OOM = 4

_L = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


class OracleBackend:
    """Constructor interface matching model3d_amd.solids.LibBackend.

    The oracle tree owns its children, so composite constructors consume the
    child handles and ``release`` of a child is a no-op."""

    def __init__(self):
        self.L = lib()
        self._keep = []

    def sphere(self, c, r):
        return self.L.or_sphere(_D3(*c), r)

    def rect(self, mn, mx):
        return self.L.or_rect(_D3(*mn), _D3(*mx))

    def capsule(self, p1, p2, r):
        return self.L.or_capsule(_D3(*p1), _D3(*p2), r)

    def cylinder(self, p1, p2, r):
        return self.L.or_cylinder(_D3(*p1), _D3(*p2), r)

    def cone(self, tip, base, r):
        return self.L.or_cone(_D3(*tip), _D3(*base), r)

    def cone_slice(self, p1, p2, r1, r2):
        return self.L.or_cone_slice(_D3(*p1), _D3(*p2), r1, r2)

    def torus(self, c, axis, outer, inner):
        return self.L.or_torus(_D3(*c), _D3(*axis), outer, inner)

    def joined(self, hs):
        return self.L.or_joined((_P * len(hs))(*hs), len(hs))

    def intersected(self, hs):
        return self.L.or_intersected((_P * len(hs))(*hs), len(hs))

    def subtracted(self, p, n):
        return self.L.or_subtracted(p, n)

    def transform(self, c, ops):
        kinds = (ctypes.c_int * max(1, len(ops)))(*[k for k, _ in ops])
        params = (ctypes.c_double * max(9, 9 * len(ops)))(*[x for _, v in ops for x in v])
        return self.L.or_transform(c, kinds, params, len(ops))

    def force_bounds(self, c, mn, mx):
        return self.L.or_force_bounds(c, _D3(*mn), _D3(*mx))

    def collider(self, tris, inset=None, hollow=None):
        t = np.ascontiguousarray(tris, dtype=np.float64)
        if hollow is not None:
            return self.L.or_collider_hollow(t.ctypes.data_as(_dp), len(t), hollow)
        if inset is not None:
            return self.L.or_collider_inset(t.ctypes.data_as(_dp), len(t), inset)
        return self.L.or_collider(t.ctypes.data_as(_dp), len(t))

    def host(self, mn, mx, batch_fn):
        def cb(user, p):
            return 1 if bool(batch_fn(np.array([[p[0], p[1], p[2]]]))[0]) else 0

        fn = HOST_FN(cb)
        self._keep.append(fn)
        return self.L.or_host(_D3(*mn), _D3(*mx), fn, None)

    def release(self, h):
        pass  # owned by the parent (or freed by the caller of _build for a root)

    def close(self):
        self._keep.clear()


class OracleSolid:
    """A solid tree lowered into the oracle; frees itself."""

    def __init__(self, solid):
        self.backend = OracleBackend()
        self.h = solid._build(self.backend)

    def __del__(self):
        try:
            if self.h:
                self.backend.L.or_free(self.h)
        except Exception:
            pass


def oracle_bounds(solid):
    o = OracleSolid(solid)
    mn, mx = _D3(), _D3()
    o.backend.L.or_bounds(o.h, mn, mx)
    return tuple(mn), tuple(mx)


def oracle_contains(solid, pts, threads=1) -> np.ndarray:
    """Solid.Contains at every point; with threads > 1 the points are split
    into contiguous chunks answered concurrently (each call releases the GIL,
    and the restated Contains only reads the solid)."""
    p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(len(p), dtype=np.uint8)
    o = OracleSolid(solid)
    L = o.backend.L
    if threads <= 1 or len(p) < 2 * threads:
        L.or_contains_batch(o.h, p.ctypes.data_as(_dp), len(p), out.ctypes.data)
        return out.astype(bool)
    from concurrent.futures import ThreadPoolExecutor

    bounds = np.linspace(0, len(p), threads + 1).astype(np.int64)

    def chunk(i):
        a, b = int(bounds[i]), int(bounds[i + 1])
        L.or_contains_batch(o.h, p[a:].ctypes.data_as(_dp), b - a, out[a:].ctypes.data)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(chunk, range(threads)))
    return out.astype(bool)


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(
            {1: "invalid bounds", 2: "solid is true outside of bounds", 3: "vertex not on edge", 4: "out of memory"}.get(
                code, f"oracle error {code}"
            )
        )
        self.code = code


def oracle_search(solid, delta, iters, threads=None, interior=False):
    """MarchingCubesSearch on the CPU: returns (triangles (T, 9), info dict).
    With ``interior`` (MarchingCubesInterior) info["interior"] holds, per
    unique vertex, its searched position and interior point (V, 6)."""
    if threads is None:
        threads = os.cpu_count() or 1
    o = OracleSolid(solid)
    L = o.backend.L
    m = _P()
    t0 = time.perf_counter()
    fn = L.or_marching_cubes_interior if interior else L.or_marching_cubes_search
    rc = fn(o.h, float(delta), int(iters), int(threads), ctypes.byref(m))
    wall = time.perf_counter() - t0
    if rc:
        raise OracleError(rc)
    try:
        n = L.or_mesh_ntris(m)
        tris = np.ctypeslib.as_array(L.or_mesh_tris(m), shape=(n * 9,)).reshape(n, 9).copy() if n else np.zeros(
            (0, 9)
        )
        c, v, e = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        s1, s2 = ctypes.c_double(), ctypes.c_double()
        L.or_mesh_info(m, c, v, e, s1, s2)
        info = dict(
            corners=c.value, vertices=v.value, search_evals=e.value, ms_scan=s1.value, ms_search=s2.value, wall_s=wall,
            threads=threads,
        )
        if interior:
            ni = ctypes.c_long()
            p = L.or_mesh_interior(m, ctypes.byref(ni))
            info["interior"] = (
                np.ctypeslib.as_array(p, shape=(ni.value * 6,)).reshape(-1, 6).copy() if ni.value else np.zeros((0, 6))
            )
        return tris, info
    finally:
        L.or_mesh_free(m)


def oracle_inverse_apply(transform, pts) -> np.ndarray:
    """transform.Inverse().Apply at every point (rows of 3), on the CPU."""
    ops = transform._ops()
    kinds = (ctypes.c_int * max(1, len(ops)))(*[k for k, _ in ops])
    params = (ctypes.c_double * max(9, 9 * len(ops)))(*[x for _, v in ops for x in v])
    out = np.ascontiguousarray(np.array(pts, dtype=np.float64).reshape(-1, 3))
    lib().or_inverse_apply(kinds, params, len(ops), out.ctypes.data_as(_dp), len(out))
    return out


def oracle_conj(solid, delta, iters, *xforms, threads=None):
    """MarchingCubesConj (mc.go:66-71) on the CPU: the search of
    TransformSolid(JoinedTransform(xforms), solid), each vertex then mapped by
    the joined transform's inverse.  Returns triangles (T, 9)."""
    from model3d_amd.solids import JoinedTransform, TransformSolid

    joined = JoinedTransform(xforms)
    tris, info = oracle_search(TransformSolid(joined, solid), delta, iters, threads)
    return oracle_inverse_apply(joined, tris).reshape(-1, 9), info


def oracle_encode_stl(tris) -> bytes:
    """EncodeSTL (export.go:20-66) of a triangle list (T, 9) on the CPU."""
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
    out = np.zeros(84 + 50 * len(t), dtype=np.uint8)
    lib().or_encode_stl(t.ctypes.data_as(_dp), len(t), out.ctypes.data)
    return out.tobytes()


def oracle_search_filter(solid, rect_filter, delta, iters, threads=None):
    """MarchingCubesSearchFilter (MarchingCubesFilter for iters 0) on the CPU:
    rect_filter(min, max) -> bool for each block.  Returns triangles (T, 9)."""
    if threads is None:
        threads = os.cpu_count() or 1
    o = OracleSolid(solid)
    L = o.backend.L

    def cb(user, mn, mx):
        return 1 if rect_filter((mn[0], mn[1], mn[2]), (mx[0], mx[1], mx[2])) else 0

    fn = RECT_FN(cb)
    m = _P()
    rc = L.or_marching_cubes_filter(o.h, fn, None, float(delta), int(iters), int(threads), ctypes.byref(m))
    if rc:
        raise OracleError(rc)
    try:
        n = L.or_mesh_ntris(m)
        return np.ctypeslib.as_array(L.or_mesh_tris(m), shape=(n * 9,)).reshape(n, 9).copy() if n else np.zeros((0, 9))
    finally:
        L.or_mesh_free(m)


def oracle_rect_collider(tris):
    """MeshToCollider(mesh).RectCollision for a triangle list (T, 9): a
    function of (min, max), brute force over the triangles."""
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
    L = lib()

    def collides(mn, mx):
        return bool(L.or_tris_rect_collision(t.ctypes.data_as(_dp), len(t), _D3(*mn), _D3(*mx)))

    return collides


def oracle_tri_sphere(tri, center, r) -> bool:
    """Triangle.SphereCollision (primitives.go:251-279) of one triangle (9
    doubles) and a sphere."""
    t = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(9))
    return bool(lib().or_tri_sphere_collision(t.ctypes.data_as(_dp), _D3(*center), float(r)))


def oracle_group_triangles(tris) -> np.ndarray:
    """GroupTriangles (bvh.go:118-156, 219-255) of a triangle list (T, 9):
    the grouped order as indices into it (sort ties broken by index)."""
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9))
    perm = np.zeros(len(t), dtype=np.uint32)
    rc = lib().or_group_triangles(t.ctypes.data_as(_dp), len(t), perm.ctypes.data)
    if rc:
        raise OracleError(rc)
    return perm


def oracle_c2f(solid, big_delta, small_delta, extra_space, iters, threads=None):
    """MarchingCubesC2F (mc.go:145-180) on the CPU."""
    import math

    extra = extra_space + 2 * big_delta * math.sqrt(3)
    coarse, _ = oracle_search(solid, big_delta, iters, threads)
    collides = oracle_rect_collider(coarse)

    def f(mn, mx):  # Rect.Expand, shapes.go:339-344
        return collides(tuple(v + -extra for v in mn), tuple(v + extra for v in mx))

    return oracle_search_filter(solid, f, small_delta, iters, threads)


def oracle_case_table() -> np.ndarray:
    out = np.zeros(256 * 31, dtype=np.int32)
    conflicts = lib().or_case_table(out.ctypes.data)
    return out.reshape(256, 31), conflicts


def oracle_rotations() -> np.ndarray:
    out = np.zeros(24 * 8, dtype=np.int32)
    n = lib().or_rotations(out.ctypes.data)
    return out.reshape(24, 8)[:n]


def oracle_spacer(solid, delta):
    o = OracleSolid(solid)
    axes = []
    for a in range(3):
        n = o.backend.L.or_spacer(o.h, float(delta), a, None)
        buf = np.zeros(n, dtype=np.float64)
        o.backend.L.or_spacer(o.h, float(delta), a, buf.ctypes.data)
        axes.append(buf)
    return axes
