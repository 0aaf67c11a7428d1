"""Python mirror of model3d's Solid API (the plugin interface of the path).

Every class here keeps the reference's name, fields and Min/Max semantics so
that code written against model3d reads the same.  A solid is lowered to the
library through ``_build(backend)``; the library evaluates Contains on the
GPU.  A user subclass of :class:`Solid` that only defines ``Min``, ``Max``
and ``Contains`` (or ``ContainsBatch``) is evaluated on the host in batches
and fed to the same GPU stages (the reference's opaque closures).

Reference: model3d/solid.go, shapes.go, transform.go, matrix.go.
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

from . import _lib


def _c3(v) -> tuple:
    t = tuple(float(x) for x in v)
    if len(t) != 3:
        raise ValueError("expected a 3-vector")
    return t


def XYZ(x: float, y: float, z: float) -> tuple:
    """model3d.XYZ (coords.go:171-173)."""
    return (float(x), float(y), float(z))


# ---------------------------------------------------------------------------
# Solid interface
# ---------------------------------------------------------------------------
class Solid:
    """model3d.Solid (solid.go:16-22): Min, Max and Contains.

    Subclasses the library knows are lowered to the GPU.  Any other subclass
    must implement Min(), Max() and Contains(p) (or ContainsBatch(pts) for an
    (n, 3) float64 array, returning n booleans); it is then evaluated on the
    host in batches.
    """

    def Min(self) -> tuple:
        return _bounds(self)[0]

    def Max(self) -> tuple:
        return _bounds(self)[1]

    def Contains(self, p) -> bool:  # pragma: no cover - user solids override
        raise NotImplementedError

    def ContainsBatch(self, pts: np.ndarray) -> np.ndarray:
        return np.fromiter((bool(self.Contains(tuple(p))) for p in pts), dtype=bool, count=len(pts))

    def _build(self, b):
        """Host-evaluated fallback for solids the library cannot lower."""
        mn, mx = _c3(self.Min()), _c3(self.Max())
        return b.host(mn, mx, self.ContainsBatch)


def _bounds(s: Solid):
    """Solid.Min/Max computed by the library with the reference's float64 ops."""
    b = LibBackend()
    h = s._build(b)
    try:
        mn, mx = _lib._D3(), _lib._D3()
        _lib.check(_lib.lib().mc3d_solid_bounds(h, mn, mx), "mc3d_solid_bounds")
        return tuple(mn), tuple(mx)
    finally:
        b.release(h)
        b.close()


class Sphere(Solid):
    """model3d.Sphere (shapes.go:13-31)."""

    def __init__(self, Center=(0.0, 0.0, 0.0), Radius: float = 1.0):
        self.Center, self.Radius = _c3(Center), float(Radius)

    def _build(self, b):
        return b.sphere(self.Center, self.Radius)


class Rect(Solid):
    """model3d.Rect (shapes.go:145-173)."""

    def __init__(self, MinVal=(0.0, 0.0, 0.0), MaxVal=(1.0, 1.0, 1.0)):
        self.MinVal, self.MaxVal = _c3(MinVal), _c3(MaxVal)

    def _build(self, b):
        return b.rect(self.MinVal, self.MaxVal)


def NewRect(min, max) -> Rect:
    """model3d.NewRect (shapes.go:151-153)."""
    return Rect(min, max)


class Capsule(Solid):
    """model3d.Capsule (shapes.go:348-368)."""

    def __init__(self, P1, P2, Radius: float):
        self.P1, self.P2, self.Radius = _c3(P1), _c3(P2), float(Radius)

    def _build(self, b):
        return b.capsule(self.P1, self.P2, self.Radius)


class Cylinder(Solid):
    """model3d.Cylinder (shapes.go:542-597)."""

    def __init__(self, P1, P2, Radius: float):
        self.P1, self.P2, self.Radius = _c3(P1), _c3(P2), float(Radius)

    def _build(self, b):
        return b.cylinder(self.P1, self.P2, self.Radius)


class Cone(Solid):
    """model3d.Cone (shapes.go:859-898)."""

    def __init__(self, Tip, Base, Radius: float):
        self.Tip, self.Base, self.Radius = _c3(Tip), _c3(Base), float(Radius)

    def _build(self, b):
        return b.cone(self.Tip, self.Base, self.Radius)


class ConeSlice(Solid):
    """model3d.ConeSlice (shapes.go:1039-1084)."""

    def __init__(self, P1, P2, R1: float, R2: float):
        self.P1, self.P2, self.R1, self.R2 = _c3(P1), _c3(P2), float(R1), float(R2)

    def _build(self, b):
        return b.cone_slice(self.P1, self.P2, self.R1, self.R2)


class Torus(Solid):
    """model3d.Torus (shapes.go:1333-1365)."""

    def __init__(self, Center, Axis, OuterRadius: float, InnerRadius: float):
        self.Center, self.Axis = _c3(Center), _c3(Axis)
        self.OuterRadius, self.InnerRadius = float(OuterRadius), float(InnerRadius)

    def _build(self, b):
        return b.torus(self.Center, self.Axis, self.OuterRadius, self.InnerRadius)


class JoinedSolid(Solid):
    """model3d.JoinedSolid (solid.go:70-96): union, bounds = union of bounds."""

    def __init__(self, solids: Iterable[Solid]):
        self.solids = list(solids)
        if not self.solids:
            raise ValueError("JoinedSolid needs at least one solid")

    def _build(self, b):
        hs = [s._build(b) for s in self.solids]
        try:
            return b.joined(hs)
        finally:
            for h in hs:
                b.release(h)


class IntersectedSolid(Solid):
    """model3d.IntersectedSolid (solid.go:142-170)."""

    def __init__(self, solids: Iterable[Solid]):
        self.solids = list(solids)
        if not self.solids:
            raise ValueError("IntersectedSolid needs at least one solid")

    def _build(self, b):
        hs = [s._build(b) for s in self.solids]
        try:
            return b.intersected(hs)
        finally:
            for h in hs:
                b.release(h)


class SubtractedSolid(Solid):
    """model3d.SubtractedSolid (solid.go:118-140)."""

    def __init__(self, Positive: Solid, Negative: Solid):
        self.Positive, self.Negative = Positive, Negative

    def _build(self, b):
        p, n = self.Positive._build(b), self.Negative._build(b)
        try:
            return b.subtracted(p, n)
        finally:
            b.release(p)
            b.release(n)


def Subtract(x: Solid, y: Solid) -> SubtractedSolid:
    """model3d.Subtract (solid.go:126-128)."""
    return SubtractedSolid(x, y)


def ClipSolid(s: Solid, min, max) -> Solid:
    """model3d.ClipSolid (solid.go:232-234)."""
    return IntersectedSolid([s, NewRect(min, max)])


class ForceSolidBounds(Solid):
    """model3d.ForceSolidBounds (solid.go:307-309)."""

    def __init__(self, s: Solid, min, max):
        self.s, self.min, self.max = s, _c3(min), _c3(max)

    def _build(self, b):
        c = self.s._build(b)
        try:
            return b.force_bounds(c, self.min, self.max)
        finally:
            b.release(c)


def CacheSolidBounds(s: Solid) -> Solid:
    """model3d.CacheSolidBounds (solid.go:316-318)."""
    return ForceSolidBounds(s, s.Min(), s.Max())


# ---------------------------------------------------------------------------
# Transforms (transform.go, matrix.go)
# ---------------------------------------------------------------------------
class Transform:
    def _ops(self) -> list:
        raise NotImplementedError


class Translate(Transform):
    """model3d.Translate (transform.go:38-58)."""

    def __init__(self, Offset):
        self.Offset = _c3(Offset)

    def _ops(self):
        return [(_lib.MC3D_XF_TRANSLATE, self.Offset + (0.0,) * 6)]


class Scale(Transform):
    """model3d.Scale (transform.go:64-93)."""

    def __init__(self, Scale: float):
        self.Scale = float(Scale)

    def _ops(self):
        return [(_lib.MC3D_XF_SCALE, (self.Scale,) + (0.0,) * 8)]


class VecScale(Transform):
    """model3d.VecScale (transform.go:95-122)."""

    def __init__(self, Scale):
        self.Scale = _c3(Scale)

    def _ops(self):
        return [(_lib.MC3D_XF_VECSCALE, self.Scale + (0.0,) * 6)]


class Matrix3Transform(Transform):
    """model3d.Matrix3Transform (transform.go:124-155); Matrix row-major."""

    def __init__(self, Matrix: Sequence[float]):
        m = tuple(float(x) for x in np.asarray(Matrix, dtype=np.float64).reshape(9))
        self.Matrix = m

    def _ops(self):
        return [(_lib.MC3D_XF_MATRIX, self.Matrix)]


class JoinedTransform(Transform):
    """model3d.JoinedTransform (transform.go:191-219), applied left to right."""

    def __init__(self, transforms: Iterable[Transform]):
        self.transforms = list(transforms)

    def _ops(self):
        out = []
        for t in self.transforms:
            out.extend(t._ops())
        return out


def _ortho_basis(c):
    """Coord3D.OrthoBasis (coords.go:388-421), same float64 operation order."""
    cx, cy, cz = c
    ax, ay, az = abs(cx), abs(cy), abs(cz)
    b1 = [0.0, 0.0, 0.0]
    if ax > ay and ax > az:
        b1[0] = cy / ax
        b1[1] = -cx / ax
    else:
        b1[1] = cz
        b1[2] = -cy
        d = ay if ay > az else az
        b1[1] /= d
        b1[2] /= d
    b2 = [b1[1] * cz - b1[2] * cy, b1[2] * cx - b1[0] * cz, b1[0] * cy - b1[1] * cx]

    def norm(v):
        n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        s = 1 / n
        return (v[0] * s, v[1] * s, v[2] * s)

    return norm(b1), norm(b2)


def _mat_mul(m, n):
    """Matrix3.Mul (matrix.go:105-121)."""
    return tuple(
        m[3 * r] * n[c] + m[3 * r + 1] * n[3 + c] + m[3 * r + 2] * n[6 + c] for r in range(3) for c in range(3)
    )


def NewMatrix3Rotation(axis, angle: float) -> tuple:
    """model3d.NewMatrix3Rotation (matrix.go:25-40).

    Uses the host libm cos/sin, which may differ from Go's math.Cos/Sin in
    the last bit; pass an explicit Matrix3Transform to reproduce a Go-built
    matrix exactly.
    """
    axis = _c3(axis)
    x, y = _ortho_basis(axis)
    basis = (axis[0], x[0], y[0], axis[1], x[1], y[1], axis[2], x[2], y[2])
    c, s = math.cos(angle), math.sin(angle)
    rot = (1.0, 0.0, 0.0, 0.0, c, s, 0.0, -s, c)
    bt = (basis[0], basis[3], basis[6], basis[1], basis[4], basis[7], basis[2], basis[5], basis[8])
    return _mat_mul(_mat_mul(basis, rot), bt)


def Rotation(axis, theta: float) -> Matrix3Transform:
    """model3d.Rotation (transform.go:162-171)."""
    return Matrix3Transform(NewMatrix3Rotation(axis, theta))


def Mirror(axis) -> Matrix3Transform:
    """model3d.Mirror (transform.go:173-185)."""
    ax = _c3(axis)
    n = math.sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2])
    s = 1 / n
    d = (ax[0] * s, ax[1] * s, ax[2] * s)
    outer = [d[i] * d[j] * -2 for i in range(3) for j in range(3)]
    ident = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    return Matrix3Transform(tuple(ident[i] + outer[i] for i in range(9)))


class TransformedSolid(Solid):
    """TransformSolid(t, s) (transform.go:240-246): bounds-checked, evaluated
    at t.Inverse().Apply(c)."""

    def __init__(self, t: Transform, s: Solid):
        self.t, self.s = t, s

    def _build(self, b):
        c = self.s._build(b)
        try:
            return b.transform(c, self.t._ops())
        finally:
            b.release(c)


def TransformSolid(t: Transform, s: Solid) -> Solid:
    return TransformedSolid(t, s)


def TranslateSolid(s: Solid, offset) -> Solid:
    """model3d.TranslateSolid (transform.go:60-62)."""
    return TransformedSolid(Translate(offset), s)


def ScaleSolid(s: Solid, factor: float) -> Solid:
    """model3d.ScaleSolid (transform.go:89-93)."""
    return TransformedSolid(Scale(factor), s)


def VecScaleSolid(s: Solid, v) -> Solid:
    """model3d.VecScaleSolid (transform.go:118-122)."""
    return TransformedSolid(VecScale(v), s)


def RotateSolid(s: Solid, axis, angle: float) -> Solid:
    """model3d.RotateSolid (transform.go:233-236)."""
    return TransformedSolid(Rotation(axis, angle), s)


# ---------------------------------------------------------------------------
# Mesh colliders
# ---------------------------------------------------------------------------
class MeshCollider:
    """MeshToCollider(mesh) (collisions.go:138-142): the triangles of a mesh."""

    def __init__(self, triangles: np.ndarray):
        t = np.ascontiguousarray(np.asarray(triangles, dtype=np.float64).reshape(-1, 9))
        if len(t) == 0:
            raise ValueError("empty mesh")
        self.triangles = t
        self._rect = None

    def _rect_handle(self):
        """The library's rect collider over these triangles (built once)."""
        if self._rect is None:
            L = _lib.lib()
            h = L.mc3d_rect_collider_create(self.triangles.ctypes.data_as(_lib._dp), len(self.triangles))
            if not h:
                raise _lib.MC3DError(_lib.MC3D_E_OOM, _lib.last_error())
            self._rect = h
        return self._rect

    def RectCollision(self, rect) -> bool:
        """MultiCollider.RectCollision (collisions.go:344-356,
        primitives.go:448-495): does any triangle touch the rect."""
        L = _lib.lib()
        return bool(L.mc3d_rect_collider_collides(self._rect_handle(), _lib._D3(*rect.MinVal), _lib._D3(*rect.MaxVal)))

    def __del__(self):
        if getattr(self, "_rect", None):
            try:
                _lib.lib().mc3d_rect_collider_free(self._rect)
            except Exception:
                pass
            self._rect = None


def MeshToCollider(mesh) -> MeshCollider:
    tris = mesh.triangles() if hasattr(mesh, "triangles") else mesh
    return MeshCollider(tris)


class ColliderSolid(Solid):
    """NewColliderSolid(c) (solid.go:247-300): ray-parity containment; with an
    inset, ColliderContains(c, p, inset) (collisions.go:119-134); with a
    radius, the hollow shell SphereCollision(p, radius)."""

    def __init__(self, collider: MeshCollider, inset: float = 0.0, radius: float = 0.0):
        self.collider = collider
        self.inset = float(inset)
        self.radius = float(radius)

    def _build(self, b):
        if self.radius != 0:
            return b.collider(self.collider.triangles, hollow=self.radius)
        if self.inset != 0:
            return b.collider(self.collider.triangles, inset=self.inset)
        return b.collider(self.collider.triangles)


def NewColliderSolid(c: MeshCollider) -> ColliderSolid:
    return ColliderSolid(c)


def NewColliderSolidInset(c: MeshCollider, inset: float) -> ColliderSolid:
    """NewColliderSolidInset (solid.go:260-270): containment at least `inset`
    from the surface (negative: also within -inset outside it)."""
    return ColliderSolid(c, inset=inset)


def NewColliderSolidHollow(c: MeshCollider, r: float) -> ColliderSolid:
    """NewColliderSolidHollow (solid.go:272-279): the shell of points within
    r of the surface."""
    return ColliderSolid(c, radius=r)


# ---------------------------------------------------------------------------
# Library backend: lowers a solid tree through the C-ABI constructors.
# ---------------------------------------------------------------------------
class LibBackend:
    """Builds mc3d_solid handles.  Host-evaluated solids keep their ctypes
    callbacks alive in this object, so keep it alive while the handle is used."""

    def __init__(self):
        self.L = _lib.lib()
        self._keep = []

    def _chk(self, h, what):
        if not h:
            raise _lib.MC3DError(_lib.MC3D_E_INVALID_ARG, f"{what}: {_lib.last_error()}")
        return h

    def sphere(self, c, r):
        return self._chk(self.L.mc3d_sphere(_lib._D3(*c), r), "mc3d_sphere")

    def rect(self, mn, mx):
        return self._chk(self.L.mc3d_rect(_lib._D3(*mn), _lib._D3(*mx)), "mc3d_rect")

    def capsule(self, p1, p2, r):
        return self._chk(self.L.mc3d_capsule(_lib._D3(*p1), _lib._D3(*p2), r), "mc3d_capsule")

    def cylinder(self, p1, p2, r):
        return self._chk(self.L.mc3d_cylinder(_lib._D3(*p1), _lib._D3(*p2), r), "mc3d_cylinder")

    def cone(self, tip, base, r):
        return self._chk(self.L.mc3d_cone(_lib._D3(*tip), _lib._D3(*base), r), "mc3d_cone")

    def cone_slice(self, p1, p2, r1, r2):
        return self._chk(self.L.mc3d_cone_slice(_lib._D3(*p1), _lib._D3(*p2), r1, r2), "mc3d_cone_slice")

    def torus(self, c, axis, outer, inner):
        return self._chk(self.L.mc3d_torus(_lib._D3(*c), _lib._D3(*axis), outer, inner), "mc3d_torus")

    def _arr(self, hs):
        return (_lib._P * len(hs))(*hs)

    def joined(self, hs):
        return self._chk(self.L.mc3d_joined(self._arr(hs), len(hs)), "mc3d_joined")

    def intersected(self, hs):
        return self._chk(self.L.mc3d_intersected(self._arr(hs), len(hs)), "mc3d_intersected")

    def subtracted(self, p, n):
        return self._chk(self.L.mc3d_subtracted(p, n), "mc3d_subtracted")

    def transform(self, c, ops):
        arr = (_lib.mc3d_xform * max(1, len(ops)))()
        for i, (k, v) in enumerate(ops):
            arr[i].kind = k
            for j in range(9):
                arr[i].v[j] = v[j]
        return self._chk(self.L.mc3d_transform(c, arr, len(ops)), "mc3d_transform")

    def force_bounds(self, c, mn, mx):
        return self._chk(self.L.mc3d_force_bounds(c, _lib._D3(*mn), _lib._D3(*mx)), "mc3d_force_bounds")

    def collider(self, tris, inset=None, hollow=None):
        t = np.ascontiguousarray(tris, dtype=np.float64)
        if hollow is not None:
            return self._chk(self.L.mc3d_collider_solid_hollow(t.ctypes.data_as(_lib._dp), len(t), hollow),
                             "mc3d_collider_solid_hollow")
        if inset is not None:
            return self._chk(self.L.mc3d_collider_solid_inset(t.ctypes.data_as(_lib._dp), len(t), inset),
                             "mc3d_collider_solid_inset")
        return self._chk(
            self.L.mc3d_collider_solid(t.ctypes.data_as(_lib._dp), len(t)), "mc3d_collider_solid"
        )

    def host(self, mn, mx, batch_fn):
        def cb(user, pts, n, out):
            try:
                if n:
                    p = np.ctypeslib.as_array(pts, shape=(n * 3,)).reshape(n, 3)
                    r = np.asarray(batch_fn(p.copy()), dtype=bool)
                    o = np.ctypeslib.as_array(out, shape=(n,))
                    o[:] = r.astype(np.uint8)
                return 0
            except Exception:  # an exception must not cross the C boundary
                return 1

        fn = _lib.CONTAINS_FN(cb)
        self._keep.append(fn)
        return self._chk(self.L.mc3d_host_solid(_lib._D3(*mn), _lib._D3(*mx), fn, None), "mc3d_host_solid")

    def release(self, h):
        if h:
            self.L.mc3d_solid_release(h)

    def close(self):
        self._keep.clear()
