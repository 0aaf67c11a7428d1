"""ctypes binding of libmc3d.so (the C-ABI declared in include/mc3d.h).

The shared library is built in place by ``model3d_amd/csrc/Makefile`` (or
``__graft_entry__.build()``).  Nothing here falls back to a CPU path: if the
library is missing or has no usable GPU, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmc3d.so")

MC3D_OK = 0
MC3D_E_INVALID_ARG = 1
MC3D_E_INVALID_BOUNDS = 2
MC3D_E_INSIDE_BOUNDARY = 3
MC3D_E_NOT_ON_EDGE = 4
MC3D_E_HIP = 5
MC3D_E_OOM = 6
MC3D_E_UNSUPPORTED = 7
MC3D_E_CALLBACK = 8
MC3D_E_INTERNAL = 9
MC3D_E_CAPACITY = 10

MC3D_XF_TRANSLATE = 1
MC3D_XF_SCALE = 2
MC3D_XF_VECSCALE = 3
MC3D_XF_MATRIX = 4

MC3D_PREC_FP64 = 0
MC3D_PREC_FP32_GUARD = 1

_D3 = ctypes.c_double * 3
_P = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class mc3d_xform(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("v", ctypes.c_double * 9)]


class mc3d_options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("cull", ctypes.c_int32),
        ("device_only", ctypes.c_int32),
        ("count_mismatch", ctypes.c_int32),
        ("brick", ctypes.c_int32),
        ("rows_per_walk", ctypes.c_int32),
        ("waves_per_brick", ctypes.c_int32),
        ("no_graph", ctypes.c_int32),
        ("slab_planes", ctypes.c_int32),
    ]


class mc3d_stats(ctypes.Structure):
    _fields_ = [
        ("corners", ctypes.c_uint64),
        ("vertices", ctypes.c_uint64),
        ("triangles", ctypes.c_uint64),
        ("bisection_evals", ctypes.c_uint64),
        ("guard_evals", ctypes.c_uint64),
        ("guard_rechecks", ctypes.c_uint64),
        ("guard_mismatches", ctypes.c_uint64),
        ("guard_violations", ctypes.c_uint64),
        ("bricks_total", ctypes.c_uint64),
        ("bricks_active", ctypes.c_uint64),
        ("corner_flops64", ctypes.c_uint64),
        ("corner_flops32", ctypes.c_uint64),
        ("bisect_flops64", ctypes.c_uint64),
        ("bisect_flops32", ctypes.c_uint64),
        ("grid_nx", ctypes.c_uint64),
        ("grid_ny", ctypes.c_uint64),
        ("grid_nz", ctypes.c_uint64),
        ("brick_code_total", ctypes.c_uint64),
        ("brick_code_max", ctypes.c_uint64),
        ("bricks_full", ctypes.c_uint64),
        ("program_length", ctypes.c_uint64),
        ("ms_device", ctypes.c_double),
        ("ms_plan", ctypes.c_double),
        ("ms_fill", ctypes.c_double),
        ("ms_corners", ctypes.c_double),
        ("ms_classify", ctypes.c_double),
        ("ms_scan", ctypes.c_double),
        ("ms_emit", ctypes.c_double),
        ("ms_bisect", ctypes.c_double),
        ("ms_d2h", ctypes.c_double),
        ("ms_host_callbacks", ctypes.c_double),
        ("ms_call", ctypes.c_double),
        ("ms_prepare", ctypes.c_double),
        ("ms_enqueue", ctypes.c_double),
        ("ms_wait", ctypes.c_double),
        ("words_listed", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


CONTAINS_FN = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, _dp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint8)
)
RECT_FILTER_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, _dp)

# Every symbol include/mc3d.h declares, with (restype, argtypes).
SIGNATURES = {
    "mc3d_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "mc3d_ctx_destroy": (None, [_P]),
    "mc3d_sphere": (_P, [_D3, ctypes.c_double]),
    "mc3d_rect": (_P, [_D3, _D3]),
    "mc3d_capsule": (_P, [_D3, _D3, ctypes.c_double]),
    "mc3d_cylinder": (_P, [_D3, _D3, ctypes.c_double]),
    "mc3d_cone": (_P, [_D3, _D3, ctypes.c_double]),
    "mc3d_cone_slice": (_P, [_D3, _D3, ctypes.c_double, ctypes.c_double]),
    "mc3d_torus": (_P, [_D3, _D3, ctypes.c_double, ctypes.c_double]),
    "mc3d_joined": (_P, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "mc3d_subtracted": (_P, [_P, _P]),
    "mc3d_intersected": (_P, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "mc3d_transform": (_P, [_P, ctypes.POINTER(mc3d_xform), ctypes.c_size_t]),
    "mc3d_force_bounds": (_P, [_P, _D3, _D3]),
    "mc3d_collider_solid": (_P, [_dp, ctypes.c_size_t]),
    "mc3d_collider_solid_inset": (_P, [_dp, ctypes.c_size_t, ctypes.c_double]),
    "mc3d_collider_solid_hollow": (_P, [_dp, ctypes.c_size_t, ctypes.c_double]),
    "mc3d_host_solid": (_P, [_D3, _D3, CONTAINS_FN, ctypes.c_void_p]),
    "mc3d_solid_retain": (_P, [_P]),
    "mc3d_solid_release": (None, [_P]),
    "mc3d_solid_bounds": (ctypes.c_int, [_P, _D3, _D3]),
    "mc3d_marching_cubes_search": (
        ctypes.c_int,
        [_P, _P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_options), ctypes.POINTER(_P)],
    ),
    "mc3d_marching_cubes_interior": (
        ctypes.c_int,
        [_P, _P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_options), ctypes.POINTER(_P)],
    ),
    "mc3d_marching_cubes_conj": (
        ctypes.c_int,
        [_P, _P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_xform), ctypes.c_size_t,
         ctypes.POINTER(mc3d_options), ctypes.POINTER(_P)],
    ),
    "mc3d_marching_cubes_search_batch": (
        ctypes.c_int,
        [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t,
         ctypes.POINTER(mc3d_options), ctypes.POINTER(_P)],
    ),
    "mc3d_mesh_counts": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "mc3d_mesh_copy": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_void_p]),
    "mc3d_mesh_copy_interior": (ctypes.c_int, [_P, ctypes.c_void_p]),
    "mc3d_mesh_host_ptrs": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
    "mc3d_marching_cubes_search_noalloc": (
        ctypes.c_int,
        [_P, _P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_options), ctypes.c_void_p, ctypes.c_size_t,
         ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    ),
    "mc3d_stl_bytes": (ctypes.c_size_t, [ctypes.c_size_t]),
    "mc3d_marching_cubes_filter": (
        ctypes.c_int,
        [_P, _P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(mc3d_options), ctypes.POINTER(_P)],
    ),
    "mc3d_marching_cubes_search_filter": (
        ctypes.c_int,
        [_P, _P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_options),
         ctypes.POINTER(_P)],
    ),
    "mc3d_marching_cubes_c2f": (
        ctypes.c_int,
        [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(mc3d_options),
         ctypes.POINTER(_P)],
    ),
    "mc3d_rect_collider_create": (_P, [_dp, ctypes.c_size_t]),
    "mc3d_rect_collider_collides": (ctypes.c_int, [_P, _D3, _D3]),
    "mc3d_rect_collider_free": (None, [_P]),
    "mc3d_mesh_encode_stl": (ctypes.c_int, [_P, _P, ctypes.c_void_p, ctypes.c_size_t]),
    "mc3d_encode_stl": (
        ctypes.c_int,
        [_P, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t],
    ),
    "mc3d_group_triangles": (
        ctypes.c_int,
        [_P, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p],
    ),
    "mc3d_mesh_group_triangles": (ctypes.c_int, [_P, _P]),
    "mc3d_mesh_device_ptrs": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
    "mc3d_mesh_stats": (ctypes.c_int, [_P, ctypes.POINTER(mc3d_stats)]),
    "mc3d_mesh_free": (None, [_P]),
    "mc3d_contains": (
        ctypes.c_int,
        [_P, _P, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_void_p],
    ),
    "mc3d_case_table": (ctypes.c_int, [ctypes.c_void_p]),
    "mc3d_diag_stream": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "mc3d_selftest_plan_fold": (
        ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    ),
    "mc3d_last_error": (ctypes.c_char_p, []),
    "mc3d_version": (ctypes.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


class MC3DError(RuntimeError):
    """A non-zero status from the library; ``code`` is the MC3D_E_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def lib():
    """Loads libmc3d.so once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `make -C model3d_amd/csrc` "
                    "or __graft_entry__.build(); there is no CPU fallback"
                )
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
        return _lib


def last_error() -> str:
    msg = lib().mc3d_last_error()
    return msg.decode() if msg else ""


def check(code: int, what: str = "") -> None:
    if code != MC3D_OK:
        raise MC3DError(code, f"{what}: {last_error()}" if what else last_error())


_ctx = {}


def context(device: int = 0):
    """A per-thread, per-device mc3d_ctx (contexts are not thread-safe)."""
    key = (threading.get_ident(), device)
    c = _ctx.get(key)
    if c is None:
        h = ctypes.c_void_p()
        check(lib().mc3d_ctx_create(device, ctypes.byref(h)), "mc3d_ctx_create")
        c = h.value
        _ctx[key] = c
    return c
