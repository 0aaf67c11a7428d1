"""GPU parity of the entry points beside MarchingCubesSearch that reuse its
kernels: MarchingCubesInterior (mc.go:55-60) and MarchingCubesConj
(mc.go:66-71), through the C-ABI, bit-exact against the oracle."""
import ctypes

import numpy as np
import pytest

import model3d_amd as m3
from model3d_amd import _lib
from oracle.binding import oracle_conj, oracle_contains, oracle_search
from tests.helpers import canonical, same_triangles, solid_catalogue

pytestmark = pytest.mark.gpu

FP64 = m3.MC3D_PREC_FP64
GUARD = m3.MC3D_PREC_FP32_GUARD

INTERIOR_SOLIDS = ["sphere", "csg", "torus", "rotated", "collider_boxes", "host_superellipsoid", "clipped"]


def _pairs(a):
    """Rows of float64 bit patterns, as a set of tuples."""
    return {tuple(r) for r in np.ascontiguousarray(a, dtype=np.float64).view(np.uint64).tolist()}


@pytest.mark.parametrize("name", INTERIOR_SOLIDS)
@pytest.mark.parametrize("iters", [0, 3, 8])
@pytest.mark.parametrize("precision", [FP64, GUARD])
def test_interior_parity(name, iters, precision):
    s = solid_catalogue()[name]
    delta = 0.045
    print(f"interior {name} iters={iters} precision={precision}", flush=True)
    mesh, interior = m3.MarchingCubesInterior(s, delta, iters, precision=precision)
    tris, info = oracle_search(s, delta, iters, threads=8, interior=True)
    assert same_triangles(mesh.triangles(), tris)
    assert interior.shape == mesh.vertices.shape
    got = _pairs(np.concatenate([mesh.vertices, interior], axis=1))
    assert got == _pairs(info["interior"])
    # the mesh itself is MarchingCubesSearch's
    plain = m3.MarchingCubesSearch(s, delta, iters, precision=precision)
    assert np.array_equal(canonical(plain.triangles()), canonical(mesh.triangles()))
    if name != "host_superellipsoid":
        assert oracle_contains(s, interior).all()


def test_interior_map_and_shapes():
    s = m3.Sphere((0, 0, 0), 0.7)
    mesh, interior = m3.MarchingCubesInterior(s, 0.05, 4, cull=False)
    mp = m3.interior_map(mesh, interior)
    assert len(mp) == len(mesh.vertices)
    v = tuple(mesh.vertices[10])
    assert mp[v] == tuple(interior[10])
    assert (np.linalg.norm(interior, axis=1) <= 0.7).all()


def test_copy_interior_rejects_other_meshes():
    L = _lib.lib()
    ctx = _lib.context(0)
    from model3d_amd.solids import LibBackend

    b = LibBackend()
    h = m3.Sphere((0, 0, 0), 0.5)._build(b)
    out = ctypes.c_void_p()
    try:
        _lib.check(L.mc3d_marching_cubes_search(ctx, h, 0.1, 2, None, ctypes.byref(out)), "search")
        buf = np.empty((1, 3))
        assert L.mc3d_mesh_copy_interior(out, buf.ctypes.data) == _lib.MC3D_E_INVALID_ARG
    finally:
        if out.value:
            L.mc3d_mesh_free(out)
        b.release(h)
        b.close()


CONJ_CASES = {
    "scale2": [m3.Scale(2.0)],
    "translate_scale": [m3.Translate((0.3, -0.1, 0.2)), m3.Scale(1.5)],
    "rotation": [m3.Rotation((0.0, 0.6, 0.8), 0.7)],
    "mirror_vecscale": [m3.Mirror((1.0, 1.0, 0.0)), m3.VecScale((1.0, 0.5, 2.0))],
    "none": [],
}


@pytest.mark.parametrize("case", list(CONJ_CASES))
@pytest.mark.parametrize("name", ["sphere", "csg", "torus", "nested_transforms", "collider_rotated"])
@pytest.mark.parametrize("precision", [FP64, GUARD])
def test_conj_parity(case, name, precision):
    s = solid_catalogue()[name]
    xf = CONJ_CASES[case]
    print(f"conj {name} {case} precision={precision}", flush=True)
    got = m3.MarchingCubesConj(s, 0.05, 5, *xf, precision=precision)
    want, _ = oracle_conj(s, 0.05, 5, *xf, threads=8)
    assert same_triangles(got.triangles(), want)
    if not xf:  # no transform: the search of TransformSolid(empty join, s)
        assert len(got.faces) == len(m3.MarchingCubesSearch(s, 0.05, 5, precision=precision).faces)


def test_conj_without_culling_and_rejects_bad_kind():
    s = solid_catalogue()["csg"]
    a = m3.MarchingCubesConj(s, 0.05, 3, m3.Rotation((1.0, 0.0, 0.0), 0.3), cull=False)
    b = m3.MarchingCubesConj(s, 0.05, 3, m3.Rotation((1.0, 0.0, 0.0), 0.3))
    assert np.array_equal(canonical(a.triangles()), canonical(b.triangles()))

    class Bad(m3.Translate):
        def _ops(self):
            return [(99, (0.0,) * 9)]

    with pytest.raises(_lib.MC3DError):
        m3.MarchingCubesConj(s, 0.05, 3, Bad((0, 0, 0)))
