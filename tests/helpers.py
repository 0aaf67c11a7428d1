"""Shared test helpers: canonical triangle sets and a catalogue of solids."""
from __future__ import annotations

import math

import numpy as np

import model3d_amd as m3


def canonical(tris) -> np.ndarray:
    """Triangle multiset in a canonical order: rows of 9 float64 bit patterns
    (vertex order inside a triangle is kept), sorted lexicographically."""
    a = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9)).view(np.uint64)
    if len(a) == 0:
        return a
    return a[np.lexsort(a.T[::-1])]


def triangle_checksum(tris) -> int:
    """An order-independent checksum of a triangle multiset (each triangle's
    9 float64 bit patterns, in their order, mixed into 64 bits; the sum of
    those over the triangles): for meshes too large to sort."""
    a = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 9)).view(np.uint64)
    h = np.zeros(len(a), dtype=np.uint64)
    for k in range(9):
        h = (h ^ a[:, k]) * np.uint64(0x9E3779B97F4A7C15 + 2 * k)
        h ^= h >> np.uint64(31)
    return int(h.sum(dtype=np.uint64))


def same_triangles(a, b) -> bool:
    ca, cb = canonical(a), canonical(b)
    return ca.shape == cb.shape and bool(np.array_equal(ca, cb))


def edge_topology(tris: np.ndarray):
    """(directed edge multiset ok, undirected edge count ok) for a closed,
    consistently oriented manifold: every directed edge once, and its
    reverse exactly once (Mesh.NeedsRepair, mesh_ops.go:343-380)."""
    t = canonical_vertices(tris)
    edges = {}
    for tri in t:
        for i in range(3):
            a, b = tri[i], tri[(i + 1) % 3]
            edges[(a, b)] = edges.get((a, b), 0) + 1
    directed_once = all(c == 1 for c in edges.values())
    paired = all(edges.get((b, a), 0) == 1 for (a, b) in edges)
    return directed_once, paired


def canonical_vertices(tris: np.ndarray):
    v = np.ascontiguousarray(np.asarray(tris, dtype=np.float64).reshape(-1, 3)).view(np.uint64)
    keys = [tuple(r) for r in v]
    return [keys[3 * i : 3 * i + 3] for i in range(len(keys) // 3)]


def volume(tris: np.ndarray) -> float:
    t = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    det = (
        a[:, 0] * (b[:, 1] * c[:, 2] - b[:, 2] * c[:, 1])
        - a[:, 1] * (b[:, 0] * c[:, 2] - b[:, 2] * c[:, 0])
        + a[:, 2] * (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0])
    )
    return float(np.sum(det / 6.0))


class Superellipsoid(m3.Solid):
    """A user-defined solid the library cannot lower (evaluated through the
    host callback): x^4 + y^4 + z^4 <= r^4 about a centre.  Only products,
    sums and a comparison, so numpy gives the same bits per point whatever
    the batch size."""

    def __init__(self, center, r):
        self.c = np.asarray(center, dtype=np.float64)
        self.r = float(r)
        self.r4 = self.r * self.r * self.r * self.r

    def Min(self):
        return tuple(float(v) for v in self.c - self.r)

    def Max(self):
        return tuple(float(v) for v in self.c + self.r)

    def ContainsBatch(self, pts):
        q = np.asarray(pts, dtype=np.float64) - self.c
        q = q * q
        return q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] <= self.r4

    def Contains(self, p):
        return bool(self.ContainsBatch(np.asarray([p]))[0])


def solid_catalogue():
    """Named solids covering every lowered type (small enough for the oracle
    at delta ~0.05)."""
    sph = m3.Sphere((0.1, -0.2, 0.05), 0.7)
    box = m3.Rect((-0.5, -0.4, -0.3), (0.6, 0.2, 0.45))
    cap = m3.Capsule((-0.4, -0.3, 0.1), (0.5, 0.35, -0.2), 0.3)
    cyl = m3.Cylinder((-0.3, 0.4, -0.5), (0.45, -0.3, 0.5), 0.35)
    cone = m3.Cone((0.2, 0.6, 0.1), (-0.2, -0.5, -0.1), 0.45)
    cslice = m3.ConeSlice((0.0, -0.5, 0.2), (0.1, 0.55, -0.1), 0.5, 0.2)
    tor = m3.Torus((0.05, 0.0, -0.1), (0.3, 0.8, 0.5), 0.6, 0.2)
    tor_axis = m3.Torus((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.6, 0.25)
    csg = m3.SubtractedSolid(
        m3.JoinedSolid([sph, m3.Capsule((0.4, 0.5, 0.0), (0.8, 0.6, 0.2), 0.2)]),
        m3.IntersectedSolid([box, m3.Sphere((0.3, 0.0, 0.0), 0.5)]),
    )
    rot = m3.RotateSolid(cyl, (0.0, 0.6, 0.8), 0.7)
    chain = m3.TransformSolid(
        m3.JoinedTransform([m3.Scale(1.3), m3.Translate((0.1, 0.2, -0.1)), m3.VecScale((1.0, 0.8, 1.2))]), cap
    )
    nested = m3.TranslateSolid(m3.JoinedSolid([m3.ScaleSolid(box, 0.8), m3.RotateSolid(tor, (1.0, 0.0, 0.0), 0.4)]), (0.05, -0.1, 0.0))
    mirrored = m3.TransformSolid(m3.Mirror((1.0, 1.0, 0.0)), cone)
    forced = m3.ForceSolidBounds(sph, (-0.7, -1.0, -0.7), (0.9, 0.6, 0.9))
    clipped = m3.ClipSolid(m3.JoinedSolid([sph, cyl]), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    # ColliderSolids: the two boxes of TestMarchingCubesFilter (mc_test.go:55-58),
    # faces on grid planes (grazing rays), and a meshed sphere.
    boxes = m3.NewMeshRect((-1, -1, -1), (0, 0, 0)).AddMesh(m3.NewMeshRect((0.1, 0, 0), (1, 1, 1)))
    col_boxes = m3.NewColliderSolid(m3.MeshToCollider(boxes))
    from oracle.binding import oracle_search

    sph_tris, _ = oracle_search(m3.Sphere((0.05, -0.02, 0.1), 0.7), 0.1, 4, threads=8)
    col_sph = m3.NewColliderSolid(m3.MeshCollider(sph_tris))
    col_rot = m3.RotateSolid(col_sph, (0.0, 0.6, 0.8), 0.5)
    # ColliderContains with a margin (collisions.go:119-134) and the hollow
    # shell (solid.go:260-300): an inset sphere mesh, the boxes grown by an
    # outset, and a shell around the sphere mesh.
    col_inset = m3.NewColliderSolidInset(m3.MeshCollider(sph_tris), 0.09)
    col_outset = m3.NewColliderSolidInset(m3.MeshToCollider(boxes), -0.12)
    col_hollow = m3.NewColliderSolidHollow(m3.MeshCollider(sph_tris), 0.07)
    # Host-evaluated solids: alone, and under transforms inside a CSG tree
    # beside native atoms (the callback sees points in its own frame).
    host = Superellipsoid((0.05, -0.1, 0.02), 0.75)
    host_csg = m3.SubtractedSolid(
        m3.JoinedSolid([m3.RotateSolid(m3.TranslateSolid(Superellipsoid((0, 0, 0), 0.5), (0.2, 0.0, 0.1)),
                                       (0.3, 0.4, 0.5), 0.6), cap]),
        m3.Sphere((0.3, 0.1, 0.2), 0.3),
    )
    return {
        "sphere": sph,
        "rect": box,
        "capsule": cap,
        "cylinder": cyl,
        "cone": cone,
        "cone_slice": cslice,
        "torus": tor,
        "torus_axis": tor_axis,
        "csg": csg,
        "rotated": rot,
        "transform_chain": chain,
        "nested_transforms": nested,
        "mirrored": mirrored,
        "force_bounds": forced,
        "clipped": clipped,
        "collider_boxes": col_boxes,
        "collider_sphere": col_sph,
        "collider_rotated": col_rot,
        "collider_inset": col_inset,
        "collider_outset": col_outset,
        "collider_hollow": col_hollow,
        "host_superellipsoid": host,
        "host_in_csg": host_csg,
    }


def check_grouping(tris, perm):
    """The grouped order's defining properties, node by node of its implicit
    tree: a node of n > 2 triangles is the n // 2 triangles with the
    smallest box midpoints along the axis whose two halves' box areas sum
    least (the first such axis), then the rest; a node of two is in x order.
    Sort ties are broken by the triangle's index."""
    t = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    mn, mx = t.min(axis=1), t.max(axis=1)
    mid = (mn + mx) * 0.5
    assert sorted(perm.tolist()) == list(range(len(t)))

    def area(ids):
        d = mx[ids].max(axis=0) - mn[ids].min(axis=0)
        return 2 * (d[0] * (d[1] + d[2]) + d[1] * d[2])

    def node(ids):
        n = len(ids)
        if n <= 2:
            assert list(ids) == sorted(ids, key=lambda i: (mid[i, 0], i))
            return
        h = n // 2
        best, first = None, None
        for a in range(3):
            order = sorted(ids, key=lambda i: (mid[i, a], i))
            v = area(np.array(order[:h])) + area(np.array(order[h:]))
            if best is None or v < best:
                best, first = v, set(order[:h])
        assert set(ids[:h]) == first
        node(ids[:h])
        node(ids[h:])

    node(list(perm))
