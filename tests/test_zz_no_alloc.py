"""mc3d_marching_cubes_search_noalloc against mc3d_marching_cubes_search,
through the C-ABI: the same vertices and triangles, bit for bit and in the
same order, on two calls in a row (the second on a smaller grid), outputs in
caller-owned device arrays.  The arrays come from the HIP runtime the
library itself links (hipMalloc through ctypes), so one runtime serves the
whole process."""
import ctypes
import os

import numpy as np
import pytest

import model3d_amd as m3
from model3d_amd import _lib, workloads
from model3d_amd.solids import LibBackend

pytestmark = pytest.mark.gpu


def _opts(precision):
    o = _lib.mc3d_options()
    o.precision, o.cull, o.device_only = precision, 1, 1
    return o


def _reference(L, ctx, h, delta, iters, opts):
    out = ctypes.c_void_p()
    _lib.check(L.mc3d_marching_cubes_search(ctx, h, delta, iters, ctypes.byref(opts), ctypes.byref(out)), "search")
    try:
        nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
        L.mc3d_mesh_counts(out, ctypes.byref(nv), ctypes.byref(nt))
        v = np.empty((nv.value, 3))
        f = np.empty((nt.value, 3), dtype=np.uint32)
        _lib.check(L.mc3d_mesh_copy(out, v.ctypes.data, f.ctypes.data), "copy")
        return v, f
    finally:
        L.mc3d_mesh_free(out)


class DeviceArray:
    """A caller-owned device array of n items of a numpy dtype."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            _lib.lib()  # the library's runtime is loaded first; the same file is reused
            path = "/opt/rocm/lib/libamdhip64.so.7"
            cls._hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        return cls._hip

    def __init__(self, n, dtype, fill_byte):
        self.n, self.dtype = n, np.dtype(dtype)
        self.ptr = ctypes.c_void_p()
        nbytes = ctypes.c_size_t(n * self.dtype.itemsize)
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        assert self.hip().hipMemset(self.ptr, fill_byte, nbytes) == 0

    def get(self, count):
        out = np.empty(count, dtype=self.dtype)
        if count:
            assert self.hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                        ctypes.c_size_t(count * self.dtype.itemsize), 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr.value:
            self.hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _noalloc(L, ctx, h, delta, iters, opts, verts, tris):
    """One call into the caller's device arrays."""
    nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
    rc = L.mc3d_marching_cubes_search_noalloc(ctx, h, delta, iters, ctypes.byref(opts), verts.ptr, verts.n // 3,
                                              tris.ptr, tris.n // 3, ctypes.byref(nv), ctypes.byref(nt))
    _lib.check(rc, "noalloc")
    v = verts.get(3 * nv.value).reshape(-1, 3)
    f = tris.get(3 * nt.value).reshape(-1, 3)
    return v, f


@pytest.mark.parametrize("precision", [m3.MC3D_PREC_FP64, m3.MC3D_PREC_FP32_GUARD])
def test_noalloc_matches_search_twice(precision):
    L = _lib.lib()
    ctx = _lib.context(0)
    b = LibBackend()
    big = workloads.csg64_solid(2)
    small = m3.SubtractedSolid(m3.Sphere((0.1, 0.0, 0.0), 0.7), m3.Capsule((-1, 0, 0), (1, 0.2, 0.1), 0.3))
    hs = [big._build(b), small._build(b)]
    opts = _opts(precision)
    # caller-owned outputs, large enough for both meshes (filled with NaN / 0xffffffff)
    verts = DeviceArray(3 * 400_000, np.float64, 0xFF)
    tris = DeviceArray(3 * 800_000, np.uint32, 0xFF)
    try:
        for h, delta in ((hs[0], 0.01), (hs[1], 0.03)):  # the second call on a smaller grid
            print(f"noalloc precision={precision} delta={delta}", flush=True)
            want_v, want_f = _reference(L, ctx, h, delta, 8, opts)
            got_v, got_f = _noalloc(L, ctx, h, delta, 8, opts, verts, tris)
            assert len(want_f) > 1000
            assert np.array_equal(got_v.view(np.uint64), want_v.view(np.uint64))
            assert np.array_equal(got_f, want_f)
    finally:
        verts.free()
        tris.free()
        for h in hs:
            b.release(h)
        b.close()


def test_noalloc_capacity_and_limits():
    L = _lib.lib()
    ctx = _lib.context(0)
    b = LibBackend()
    h = m3.Sphere((0, 0, 0), 1.0)._build(b)
    opts = _opts(m3.MC3D_PREC_FP64)
    verts = tris = None
    try:
        nv, nt = ctypes.c_size_t(), ctypes.c_size_t()
        print("noalloc: counts with zero capacities", flush=True)
        rc = L.mc3d_marching_cubes_search_noalloc(ctx, h, 0.02, 8, ctypes.byref(opts), None, 0, None, 0,
                                                  ctypes.byref(nv), ctypes.byref(nt))
        assert rc == _lib.MC3D_E_CAPACITY
        assert (nv.value, nt.value) == (46998, 93992)
        # arrays one triangle short: an error, and nothing written past them
        NV, NT = nv.value, nt.value
        verts = DeviceArray(3 * NV + 3, np.float64, 0)
        tris = DeviceArray(3 * NT + 3, np.uint32, 0xAB)
        print("noalloc: one triangle short", flush=True)
        rc = L.mc3d_marching_cubes_search_noalloc(ctx, h, 0.02, 8, ctypes.byref(opts), verts.ptr, NV + 1,
                                                  tris.ptr, NT - 1, ctypes.byref(nv), ctypes.byref(nt))
        assert rc == _lib.MC3D_E_CAPACITY
        tail = tris.get(3 * NT + 3)[3 * (NT - 1):]
        assert len(tail) == 6 and np.all(tail == 0xABABABAB)
        # a grid beyond the scratch array: refused before anything runs
        print("noalloc: grid beyond the scratch", flush=True)
        rc = L.mc3d_marching_cubes_search_noalloc(ctx, h, 0.002, 8, ctypes.byref(opts), verts.ptr, NV, tris.ptr, NT,
                                                  ctypes.byref(nv), ctypes.byref(nt))
        assert rc == _lib.MC3D_E_UNSUPPORTED
        assert "scratch" in _lib.last_error()
    finally:
        for arr in (verts, tris):
            if arr is not None:
                arr.free()
        b.release(h)
        b.close()
