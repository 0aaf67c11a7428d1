"""MarchingCubesSearch and its result mesh (model3d/mc.go:14-49).

The reference returns a *Mesh, a set of *Triangle.  Here a mesh is the
indexed form the GPU produces: float64 vertices (V, 3) and uint32 triangles
(T, 3); ``triangles()`` expands it to the reference's (T, 3, 3) triangle
list with vertex order unchanged.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .solids import LibBackend, Solid


class Panic(RuntimeError):
    """A reference panic (invalid bounds, inside-boundary solid, ...)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float64
    faces: np.ndarray  # (T, 3) uint32
    stats: dict = field(default_factory=dict)

    def triangles(self) -> np.ndarray:
        """(T, 3, 3) float64: triangle i = vertices[faces[i]]."""
        return self.vertices[self.faces.astype(np.int64)]

    def EncodeSTL(self, device: int = 0) -> bytes:
        """Mesh.EncodeSTL (mesh.go:652-655): binary STL of the triangles in
        this mesh's order, encoded on the GPU (see EncodeSTL)."""
        return EncodeSTL(self, device=device)

    def SaveGroupedSTL(self, path, device: int = 0) -> None:
        """Mesh.SaveGroupedSTL (mesh.go:721-743): the triangles in
        GroupTriangles order (grouped on the GPU), written as binary STL."""
        g = Mesh(self.vertices, self.faces[group_order(self, device=device).astype(np.int64)])
        WriteSTL(path, g, device=device)

    @classmethod
    def from_triangles(cls, tris) -> "Mesh":
        """Indexed mesh of a triangle list (vertices shared by exact bits)."""
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 3))
        keys = t.view(np.uint64)
        _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
        return cls(t[first], inv.reshape(-1, 3).astype(np.uint32))

    def AddMesh(self, other: "Mesh") -> "Mesh":
        """Mesh.AddMesh (mesh.go:399-402): the union of both triangle lists."""
        m = Mesh.from_triangles(np.concatenate([self.triangles(), other.triangles()]))
        self.vertices, self.faces = m.vertices, m.faces
        return self

    def NumTriangles(self) -> int:
        return int(len(self.faces))

    def Volume(self) -> float:
        """Mesh.Volume (measurements.go:18-29): |sum det(t)/6|."""
        t = self.triangles()
        a, b, c = t[:, 0], t[:, 1], t[:, 2]
        det = (
            a[:, 0] * (b[:, 1] * c[:, 2] - b[:, 2] * c[:, 1])
            - a[:, 1] * (b[:, 0] * c[:, 2] - b[:, 2] * c[:, 0])
            + a[:, 2] * (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0])
        )
        return float(abs(np.sum(det / 6.0)))


def NewMeshRect(min, max) -> Mesh:
    """model3d.NewMeshRect (mesh.go:132-162): 6 quads split by AddQuad
    (mesh.go:389-397) into triangles {p1, p2, p4}, {p2, p3, p4}."""
    mn, mx = [float(v) for v in min], [float(v) for v in max]

    def point(x, y, z):
        return (mx[0] if x else mn[0], mx[1] if y else mn[1], mx[2] if z else mn[2])

    quads = [
        (point(0, 0, 0), point(1, 0, 0), point(1, 0, 1), point(0, 0, 1)),
        (point(1, 1, 1), point(1, 1, 0), point(0, 1, 0), point(0, 1, 1)),
        (point(0, 0, 0), point(0, 0, 1), point(0, 1, 1), point(0, 1, 0)),
        (point(1, 1, 1), point(1, 0, 1), point(1, 0, 0), point(1, 1, 0)),
        (point(0, 0, 0), point(0, 1, 0), point(1, 1, 0), point(1, 0, 0)),
        (point(1, 1, 1), point(0, 1, 1), point(0, 0, 1), point(1, 0, 1)),
    ]
    tris = []
    for p1, p2, p3, p4 in quads:
        tris.append((p1, p2, p4))
        tris.append((p2, p3, p4))
    return Mesh.from_triangles(np.array(tris))


def _options(precision, cull, device_only=False, count_mismatch=False, brick=0, rows_per_walk=0, waves_per_brick=0,
             slab_planes=0):
    o = _lib.mc3d_options()
    o.slab_planes = slab_planes
    o.precision = precision
    o.cull = 1 if cull else 0
    o.device_only = 1 if device_only else 0
    o.count_mismatch = 1 if count_mismatch else 0
    o.brick = brick
    o.rows_per_walk = rows_per_walk
    o.waves_per_brick = waves_per_brick
    return o


def MarchingCubesSearch(
    s: Solid,
    delta: float,
    iters: int,
    *,
    precision: int = _lib.MC3D_PREC_FP32_GUARD,
    cull: bool = True,
    device: int = 0,
    count_mismatch: bool = False,
    brick: int = 0,
    rows_per_walk: int = 0,
    waves_per_brick: int = 0,
    slab_planes: int = 0,
) -> Mesh:
    """model3d.MarchingCubesSearch (mc.go:45-49) on the GPU.

    precision: MC3D_PREC_FP32_GUARD (default) or MC3D_PREC_FP64; cull: per-brick
    culling (brick = its y/z edge in corners, 0 = default; rows_per_walk and
    waves_per_brick shape the corner kernel); count_mismatch:
    also gather float32-vs-float64 statistics into ``Mesh.stats``;
    slab_planes: stream the grid in z slabs of that many cube layers (0: only
    past 2^32 occupancy words).  Every setting yields the same triangles."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    try:
        opts = _options(precision, cull, count_mismatch=count_mismatch, brick=brick, rows_per_walk=rows_per_walk,
                        waves_per_brick=waves_per_brick, slab_planes=slab_planes)
        _raise_for(L.mc3d_marching_cubes_search(ctx, h, float(delta), int(iters), ctypes.byref(opts),
                                                ctypes.byref(out)))
        return _fetch(L, out)
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


def MarchingCubesInterior(
    s: Solid,
    delta: float,
    iters: int,
    *,
    precision: int = _lib.MC3D_PREC_FP32_GUARD,
    cull: bool = True,
    device: int = 0,
    slab_planes: int = 0,
):
    """model3d.MarchingCubesInterior (mc.go:55-60) on the GPU: the mesh of
    MarchingCubesSearch and, per vertex, a nearby point known to be inside
    the solid (the bracket's inside end, mc.go:252-255).  The reference
    returns a CoordMap from vertex to interior point; here ``interior[i]`` is
    the point of ``mesh.vertices[i]`` (``interior_map`` builds the mapping)."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    try:
        opts = _options(precision, cull, slab_planes=slab_planes)
        _raise_for(L.mc3d_marching_cubes_interior(ctx, h, float(delta), int(iters), ctypes.byref(opts),
                                                  ctypes.byref(out)))
        mesh = _fetch(L, out)
        interior = np.empty((len(mesh.vertices), 3), dtype=np.float64)
        _lib.check(L.mc3d_mesh_copy_interior(out, interior.ctypes.data), "mc3d_mesh_copy_interior")
        return mesh, interior
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


def interior_map(mesh: Mesh, interior: np.ndarray) -> dict:
    """The CoordMap[Coord3D] of MarchingCubesInterior: vertex -> interior point."""
    return {tuple(v): tuple(p) for v, p in zip(mesh.vertices.tolist(), interior.tolist())}


def MarchingCubesConj(
    s: Solid,
    delta: float,
    iters: int,
    *xforms,
    precision: int = _lib.MC3D_PREC_FP32_GUARD,
    cull: bool = True,
    device: int = 0,
) -> Mesh:
    """model3d.MarchingCubesConj (mc.go:66-71) on the GPU: MarchingCubesSearch
    of TransformSolid(JoinedTransform(xforms), s), then every vertex mapped by
    the joined transform's Inverse() (on the device)."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    ops = [op for t in xforms for op in t._ops()]
    arr = (_lib.mc3d_xform * max(1, len(ops)))()
    for i, (k, v) in enumerate(ops):
        arr[i].kind = k
        for j in range(9):
            arr[i].v[j] = v[j]
    try:
        opts = _options(precision, cull)
        _raise_for(L.mc3d_marching_cubes_conj(ctx, h, float(delta), int(iters), arr, len(ops), ctypes.byref(opts),
                                              ctypes.byref(out)))
        return _fetch(L, out)
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


def group_order(mesh_or_triangles, device: int = 0) -> np.ndarray:
    """The GroupTriangles order (bvh.go:118-156, 219-292) of a Mesh's
    triangles or of a (T, 3, 3) / (T, 9) triangle array, computed on the GPU:
    indices into the triangles (equal box midpoints keep their index order)."""
    if isinstance(mesh_or_triangles, Mesh):
        v = np.ascontiguousarray(mesh_or_triangles.vertices, dtype=np.float64)
        f = np.ascontiguousarray(mesh_or_triangles.faces, dtype=np.uint32)
    else:
        t = np.asarray(mesh_or_triangles, dtype=np.float64).reshape(-1, 9)
        v = np.ascontiguousarray(t.reshape(-1, 3))
        f = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    L = _lib.lib()
    ctx = _lib.context(device)
    perm = np.empty(len(f), dtype=np.uint32)
    _raise_for(L.mc3d_group_triangles(ctx, v.ctypes.data, len(v), f.ctypes.data, len(f), perm.ctypes.data))
    return perm


def GroupTriangles(triangles, device: int = 0) -> np.ndarray:
    """model3d.GroupTriangles (bvh.go:118-129): the triangles, (T, 3, 3), in
    the grouped order that GroupedTrianglesToCollider and SaveGroupedSTL use."""
    t = np.asarray(triangles, dtype=np.float64).reshape(-1, 3, 3)
    return t[group_order(t, device=device).astype(np.int64)]


def EncodeSTL(mesh_or_triangles, device: int = 0) -> bytes:
    """model3d.EncodeSTL (export.go:20-26) on the GPU: 80 zero bytes, the
    triangle count, then per triangle the float32 normal (Triangle.Normal in
    float64), the float32 vertices and two zero bytes.  Accepts a Mesh or a
    (T, 3, 3) / (T, 9) triangle array."""
    if isinstance(mesh_or_triangles, Mesh):
        v = np.ascontiguousarray(mesh_or_triangles.vertices, dtype=np.float64)
        f = np.ascontiguousarray(mesh_or_triangles.faces, dtype=np.uint32)
    else:
        t = np.asarray(mesh_or_triangles, dtype=np.float64).reshape(-1, 9)
        v = np.ascontiguousarray(t.reshape(-1, 3))
        f = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    L = _lib.lib()
    ctx = _lib.context(device)
    n = L.mc3d_stl_bytes(len(f))
    out = np.empty(n, dtype=np.uint8)
    _raise_for(L.mc3d_encode_stl(ctx, v.ctypes.data, len(v), f.ctypes.data, len(f), out.ctypes.data, n))
    return out.tobytes()


def WriteSTL(w, mesh_or_triangles, device: int = 0) -> None:
    """model3d.WriteSTL (export.go:28-35): the binary STL to a path or a
    writable binary file object."""
    data = EncodeSTL(mesh_or_triangles, device=device)
    if hasattr(w, "write"):
        w.write(data)
    else:
        with open(w, "wb") as fh:
            fh.write(data)


def _rect_filter(f):
    """(C function, user pointer, keep-alive) for a block filter: a
    MeshCollider's RectCollision runs natively; any other callable f(rect)
    is called back with a Rect."""
    from .solids import MeshCollider, Rect

    if getattr(f, "__func__", None) is MeshCollider.RectCollision and isinstance(f.__self__, MeshCollider):
        L = _lib.lib()
        return ctypes.cast(L.mc3d_rect_collider_collides, ctypes.c_void_p), f.__self__._rect_handle(), f.__self__

    def cb(user, mn, mx):
        return 1 if f(Rect((mn[0], mn[1], mn[2]), (mx[0], mx[1], mx[2]))) else 0

    fn = _lib.RECT_FILTER_FN(cb)
    return ctypes.cast(fn, ctypes.c_void_p), None, fn


def MarchingCubesSearchFilter(s: Solid, f, delta: float, iters: int, *, precision: int = _lib.MC3D_PREC_FP32_GUARD,
                              cull: bool = True, device: int = 0) -> Mesh:
    """model3d.MarchingCubesSearchFilter (mc.go:137-143) on the GPU: marching
    cubes only in the blocks for which f(rect) is true (the reference's block
    recursion, mc.go:73-135, 724-797), then the vertex search."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    fn, user, keep = _rect_filter(f)
    try:
        opts = _options(precision, cull)
        _raise_for(L.mc3d_marching_cubes_search_filter(ctx, h, fn, user, float(delta), int(iters), ctypes.byref(opts),
                                                       ctypes.byref(out)))
        return _fetch(L, out)
    finally:
        del keep
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


def MarchingCubesFilter(s: Solid, f, delta: float, **opts) -> Mesh:
    """model3d.MarchingCubesFilter (mc.go:73-135): MarchingCubes in the
    blocks f accepts (no vertex search)."""
    return MarchingCubesSearchFilter(s, f, delta, 0, **opts)


def MarchingCubesC2F(s: Solid, bigDelta: float, smallDelta: float, extraSpace: float, iters: int, *,
                     precision: int = _lib.MC3D_PREC_FP32_GUARD, cull: bool = True, device: int = 0) -> Mesh:
    """model3d.MarchingCubesC2F (mc.go:145-180) on the GPU: a coarse search,
    its mesh (left on the device) as the rect collider of the block
    recursion (on the device), the filtered fine search."""
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    out = ctypes.c_void_p()
    try:
        opts = _options(precision, cull)
        _raise_for(L.mc3d_marching_cubes_c2f(ctx, h, float(bigDelta), float(smallDelta), float(extraSpace), int(iters),
                                             ctypes.byref(opts), ctypes.byref(out)))
        return _fetch(L, out)
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


def _raise_for(rc: int):
    """A reference panic as Panic (same message), anything else as MC3DError."""
    if rc != _lib.MC3D_OK:
        msg = _lib.last_error()
        if rc in (_lib.MC3D_E_INVALID_BOUNDS, _lib.MC3D_E_INSIDE_BOUNDARY, _lib.MC3D_E_NOT_ON_EDGE):
            raise Panic(rc, msg)
        raise _lib.MC3DError(rc, msg)


def _fetch(L, out) -> Mesh:
    nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt)), "mc3d_mesh_counts")
    verts = np.empty((nv.value, 3), dtype=np.float64)
    faces = np.empty((nt.value, 3), dtype=np.uint32)
    _lib.check(L.mc3d_mesh_copy(out, verts.ctypes.data, faces.ctypes.data), "mc3d_mesh_copy")
    st = _lib.mc3d_stats()
    _lib.check(L.mc3d_mesh_stats(out, ctypes.byref(st)), "mc3d_mesh_stats")
    return Mesh(verts, faces, st.as_dict())


def MarchingCubesSearchBatch(
    solids,
    deltas,
    iters,
    *,
    precision: int = _lib.MC3D_PREC_FP32_GUARD,
    cull: bool = True,
    device: int = 0,
    count_mismatch: bool = False,
    brick: int = 0,
    rows_per_walk: int = 0,
    waves_per_brick: int = 0,
) -> list:
    """MarchingCubesSearch (mc.go:45-49) of every solid in one pass of the GPU
    pipeline; ``deltas`` and ``iters`` are per solid (or one value for all).
    Returns one Mesh per solid, each identical to MarchingCubesSearch's."""
    solids = list(solids)
    n = len(solids)
    deltas = [float(d) for d in (deltas if hasattr(deltas, "__len__") else [deltas] * n)]
    iters = [int(i) for i in (iters if hasattr(iters, "__len__") else [iters] * n)]
    if len(deltas) != n or len(iters) != n:
        raise ValueError("deltas and iters need one entry per solid")
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    hs = [s._build(b) for s in solids]
    outs = (ctypes.c_void_p * max(1, n))()
    try:
        opts = _options(precision, cull, count_mismatch=count_mismatch, brick=brick, rows_per_walk=rows_per_walk,
                        waves_per_brick=waves_per_brick)
        _raise_for(L.mc3d_marching_cubes_search_batch(
            ctx, (ctypes.c_void_p * max(1, n))(*hs), (ctypes.c_double * max(1, n))(*deltas),
            (ctypes.c_int32 * max(1, n))(*iters), n, ctypes.byref(opts), outs))
        return [_fetch(L, outs[i]) for i in range(n)]
    finally:
        for i in range(n):
            if outs[i]:
                L.mc3d_mesh_free(outs[i])
        for h in hs:
            b.release(h)
        b.close()


def MarchingCubes(s: Solid, delta: float, **kw) -> Mesh:
    """model3d.MarchingCubes (mc.go:14-37): MarchingCubesSearch with iters=0."""
    return MarchingCubesSearch(s, delta, 0, **kw)


def Contains(s: Solid, pts, *, precision: int = _lib.MC3D_PREC_FP64, device: int = 0) -> np.ndarray:
    """Solid.Contains for an (n, 3) array of points, evaluated on the GPU."""
    p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(len(p), dtype=np.uint8)
    L = _lib.lib()
    ctx = _lib.context(device)
    b = LibBackend()
    h = s._build(b)
    try:
        _lib.check(L.mc3d_contains(ctx, h, p.ctypes.data, len(p), precision, out.ctypes.data), "mc3d_contains")
    finally:
        b.release(h)
        b.close()
    return out.astype(bool)


def case_table() -> np.ndarray:
    """The library's 256 x 16 case table (count, then cube-edge ids)."""
    t = np.zeros(256 * 16, dtype=np.uint8)
    _lib.check(_lib.lib().mc3d_case_table(t.ctypes.data), "mc3d_case_table")
    return t.reshape(256, 16)
