"""Pins the oracle (the CPU restatement used as the checker).

Each test checks the oracle against something independent of its own code:
the reference tests' properties (model3d/mc_test.go, measurements_test.go,
shapes_test.go, solid_test.go), closed forms, textbook predicates evaluated
with numpy, and committed fixtures.  Runs on the CPU.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import model3d_amd as m3
from model3d_amd import workloads
from oracle.binding import (
    OracleError,
    oracle_bounds,
    oracle_case_table,
    oracle_conj,
    oracle_contains,
    oracle_c2f,
    oracle_encode_stl,
    oracle_inverse_apply,
    oracle_rect_collider,
    oracle_rotations,
    oracle_search,
    oracle_search_filter,
    oracle_spacer,
)
from tests.helpers import canonical, check_grouping, edge_topology, solid_catalogue, volume

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# Case table (mc.go:285-586)
# ---------------------------------------------------------------------------
def test_rotations_group():
    rots = oracle_rotations()
    assert len(rots) == 24
    as_tuples = [tuple(r) for r in rots]
    assert as_tuples == sorted(as_tuples)  # sorted order, mc.go:339-350
    assert as_tuples[0] == tuple(range(8))
    s = set(as_tuples)
    for a in as_tuples:  # closed under Compose (mc.go:356-362)
        assert sorted(a) == list(range(8))
        for b in as_tuples:
            assert tuple(a[b[i]] for i in range(8)) in s


def test_case_table_determinism_and_classes():
    t1, conflicts = oracle_case_table()
    t2, _ = oracle_case_table()
    # TestMarchingCubesDeterminism (mc_test.go:9-25)
    assert np.array_equal(t1, t2)
    # every key reached from exactly one of the 23 rotation classes, so Go's
    # random map order over baseTriangleTable cannot change the table
    assert conflicts == 0
    assert (t1[:, 0] >= 0).all()
    assert t1[0, 0] == 0 and t1[255, 0] == 0


def test_case_table_edges_change_sign_and_cover():
    t, _ = oracle_case_table()
    for k in range(256):
        inside = [(k >> c) & 1 for c in range(8)]
        used = set()
        for tri in range(t[k, 0]):
            ids = t[k, 1 + 6 * tri : 7 + 6 * tri]
            for e in range(3):
                a, b = int(ids[2 * e]), int(ids[2 * e + 1])
                d = a ^ b
                assert d in (1, 2, 4), (k, a, b)  # a cube edge
                assert inside[a] != inside[b], (k, a, b)  # crossing the surface
                used.add((min(a, b), max(a, b)))
        crossed = {
            (a, a | d) for a in range(8) for d in (1, 2, 4) if not a & d and inside[a] != inside[a | d]
        }
        # every crossed cube edge carries a vertex: one vertex per crossed grid edge
        assert used == crossed, k


def test_case_table_fixture():
    t, _ = oracle_case_table()
    with open(os.path.join(GOLDEN, "mc_case_table.json")) as f:
        fix = json.load(f)["cases"]
    for k in range(256):
        n = int(t[k, 0])
        assert [[int(c) for c in t[k, 1 + 6 * i : 7 + 6 * i]] for i in range(n)] == fix[k]["triangles"]


def _mc_go_table():
    """The table derived from mc.go's text by tests/golden/make_case_table_from_mc_go.py."""
    with open(os.path.join(GOLDEN, "mc_case_table_mc_go.json")) as f:
        fix = json.load(f)
    assert fix["base_cases"] == 23 and fix["rotations"] == 24
    return [c["triangles"] for c in fix["cases"]]


def test_case_table_matches_mc_go_derivation():
    """The oracle's table equals the one derived from mc.go's text (base
    cases parsed as data, rotation rule restated separately), triangle by
    triangle and corner by corner."""
    t, _ = oracle_case_table()
    ref = _mc_go_table()
    for k in range(256):
        n = int(t[k, 0])
        assert [[int(c) for c in t[k, 1 + 6 * i : 7 + 6 * i]] for i in range(n)] == ref[k], k


def test_case_table_hand_derived_keys():
    """Keys whose rotation is the identity (the first in sorted order) read
    baseTriangleTable directly (mc.go:462-586): corner 0 alone, corners 0 and
    1, every corner but 0; the empty and full cubes have no triangles."""
    ref = _mc_go_table()
    assert ref[1] == [[0, 1, 0, 2, 0, 4]]
    assert ref[3] == [[0, 4, 1, 5, 0, 2], [1, 5, 1, 3, 0, 2]]
    assert ref[254] == [[0, 2, 0, 1, 0, 4]]
    assert ref[0] == [] and ref[255] == []
    t, _ = oracle_case_table()
    for k in (1, 3, 254):
        n = int(t[k, 0])
        assert [[int(c) for c in t[k, 1 + 6 * i : 7 + 6 * i]] for i in range(n)] == ref[k]


# ---------------------------------------------------------------------------
# Spacer (mc.go:594-608)
# ---------------------------------------------------------------------------
def test_spacer_matches_python_float_loop():
    with open(os.path.join(GOLDEN, "spacer_sphere_r1_d0.02.json")) as f:
        fix = json.load(f)
    xs = [struct.unpack(">d", bytes.fromhex(h))[0] for h in fix["values_hex"]]
    assert len(xs) == 102 and xs[-1] == 1.000000000000001
    ax = oracle_spacer(m3.Sphere((0, 0, 0), 1.0), 0.02)
    for a in range(3):
        assert ax[a].tolist() == xs


# ---------------------------------------------------------------------------
# Sphere closed forms (measurements_test.go:9-22)
# ---------------------------------------------------------------------------
def test_sphere_config_closed_forms():
    tris, info = oracle_search(m3.Sphere((0, 0, 0), 1.0), 0.02, 8, threads=8)
    with open(os.path.join(GOLDEN, "sphere_r1_d0.02_counts.json")) as f:
        fix = json.load(f)
    assert info["corners"] == fix["corners"]
    assert len(tris) == fix["triangles"]
    assert info["vertices"] == fix["vertices"]
    assert info["search_evals"] == fix["bisection_evals_iters8"]
    # Euler: closed genus-0 triangulation has T = 2V - 4
    assert len(tris) == 2 * info["vertices"] - 4
    assert abs(volume(tris) - 4.0 / 3.0 * math.pi) < 1e-2
    directed_once, paired = edge_topology(tris)
    assert directed_once and paired


def test_mesh_volume_random_spheres():
    rng = workloads.SplitMix64(7)
    for _ in range(3):
        c = [rng.uniform(-1, 1) for _ in range(3)]
        r = rng.uniform(0.1, 1.0)
        tris, _ = oracle_search(m3.Sphere(c, r), 0.02, 8, threads=8)
        assert abs(volume(tris) - 4.0 / 3.0 * math.pi * r**3) < 1e-2


def test_bisection_converges_on_sphere():
    """Each vertex lies on a grid edge within delta / 2^(iters+1) of the exact
    crossing of the unit sphere along that edge."""
    delta, iters = 0.05, 6
    tris, _ = oracle_search(m3.Sphere((0, 0, 0), 1.0), delta, iters, threads=8)
    ax = oracle_spacer(m3.Sphere((0, 0, 0), 1.0), delta)
    grid = [set(a.tolist()) for a in ax]
    v = np.unique(tris.reshape(-1, 3), axis=0)
    for p in v:
        on = [p[i] in grid[i] for i in range(3)]
        assert sum(on) == 2
        a = on.index(False)
        rest = sum(p[i] ** 2 for i in range(3) if i != a)
        exact = math.copysign(math.sqrt(max(0.0, 1 - rest)), p[a])
        assert abs(p[a] - exact) <= delta / 2 ** (iters + 1) + 1e-12


# ---------------------------------------------------------------------------
# Contains (shapes.go, solid.go) against textbook predicates
# ---------------------------------------------------------------------------
def _far(d, tol=1e-9):
    return np.abs(d) > tol


def _seg_dist(p, a, b):
    ab = np.asarray(b) - np.asarray(a)
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0, 1)
    return np.linalg.norm(p - (np.asarray(a) + t[:, None] * ab), axis=1)


def test_contains_textbook_predicates():
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.2, 1.2, size=(20000, 3))
    cat = solid_catalogue()
    s = cat["sphere"]
    d = np.linalg.norm(p - s.Center, axis=1) - s.Radius
    got = oracle_contains(s, p)
    assert np.array_equal(got[_far(d)], (d < 0)[_far(d)])

    b = cat["rect"]
    inside = np.all((p >= b.MinVal) & (p <= b.MaxVal), axis=1)
    assert np.array_equal(oracle_contains(b, p), inside)

    c = cat["capsule"]
    d = _seg_dist(p, np.array(c.P1), np.array(c.P2)) - c.Radius
    got = oracle_contains(c, p)
    assert np.array_equal(got[_far(d)], (d < 0)[_far(d)])

    cy = cat["cylinder"]
    p1, p2 = np.array(cy.P1), np.array(cy.P2)
    ax = p1 - p2
    L = np.linalg.norm(ax)
    t = (p - p2) @ ax / L
    radial = np.linalg.norm((p - p2) - np.outer(t / L, ax), axis=1)
    margin = np.minimum(np.minimum(np.abs(t), np.abs(t - L)), np.abs(radial - cy.Radius))
    want = (t >= 0) & (t <= L) & (radial <= cy.Radius)
    got = oracle_contains(cy, p)
    assert np.array_equal(got[_far(margin)], want[_far(margin)])

    co = cat["cone"]
    tip, base = np.array(co.Tip), np.array(co.Base)
    ax = tip - base
    L = np.linalg.norm(ax)
    t = (p - base) @ ax / L
    radial = np.linalg.norm((p - base) - np.outer(t / L, ax), axis=1)
    rmax = co.Radius * (1 - t / L)
    margin = np.minimum(np.minimum(np.abs(t), np.abs(t - L)), np.abs(radial - rmax))
    want = (t >= 0) & (t <= L) & (radial <= rmax)
    got = oracle_contains(co, p)
    assert np.array_equal(got[_far(margin)], want[_far(margin)])

    cs = cat["cone_slice"]
    p1, p2 = np.array(cs.P1), np.array(cs.P2)
    ax = p2 - p1
    L = np.linalg.norm(ax)
    t = (p - p1) @ ax / L  # distance along the axis from P1
    radial = np.linalg.norm((p - p1) - np.outer(t / L, ax), axis=1)
    r = cs.R1 * (1 - t / L) + cs.R2 * (t / L)  # radius interpolates from R1 at P1 to R2 at P2
    margin = np.minimum(np.minimum(np.abs(t), np.abs(t - L)), np.abs(radial - r))
    want = (t >= 0) & (t <= L) & (radial <= r)
    got = oracle_contains(cs, p)
    assert np.array_equal(got[_far(margin)], want[_far(margin)])
    assert want[_far(margin)].sum() > 100  # the sample reaches inside

    tor = cat["torus"]
    axis = np.array(tor.Axis) / np.linalg.norm(tor.Axis)
    q = p - tor.Center
    h = q @ axis
    rho = np.linalg.norm(q - np.outer(h, axis), axis=1)
    d = np.sqrt((rho - tor.OuterRadius) ** 2 + h**2) - tor.InnerRadius
    got = oracle_contains(tor, p)
    assert np.array_equal(got[_far(d)], (d < 0)[_far(d)])


def test_cone_containment():
    """TestConeContainment (shapes_test.go:119-136): points along the axis
    and at the rim."""
    co = m3.Cone((0, 0, 1), (0, 0, -1), 0.5)
    pts = np.array([[0, 0, 0.99], [0, 0, -0.99], [0.24, 0, 0.0], [0.26, 0, 0.0], [0, 0, 1.01], [0.49, 0, -0.99]])
    assert oracle_contains(co, pts).tolist() == [True, True, True, False, False, True]


def test_csg_identities():
    """solid_test.go:8-95: Joined = OR, Subtracted = AND NOT, Intersected = AND."""
    rng = np.random.default_rng(11)
    p = rng.uniform(-1.2, 1.2, size=(5000, 3))
    cat = solid_catalogue()
    a, b = cat["sphere"], cat["cylinder"]
    ca, cb = oracle_contains(a, p), oracle_contains(b, p)
    assert np.array_equal(oracle_contains(m3.JoinedSolid([a, b]), p), ca | cb)
    assert np.array_equal(oracle_contains(m3.SubtractedSolid(a, b), p), ca & ~cb)
    assert np.array_equal(oracle_contains(m3.IntersectedSolid([a, b]), p), ca & cb)
    # translation moves the solid: Contains(p) of the moved solid == Contains(p - off)
    off = (0.25, -0.125, 0.5)  # exactly representable, so p - off is exact
    moved = m3.TranslateSolid(a, off)
    assert np.array_equal(oracle_contains(moved, p + np.array(off)), ca)


def test_solid_bounds_property():
    """testSolidBounds (shapes_test.go:87-117): Contains is false just outside
    Min/Max, on every face."""
    rng = np.random.default_rng(3)
    for name, s in solid_catalogue().items():
        mn, mx = np.array(oracle_bounds(s)[0]), np.array(oracle_bounds(s)[1])
        for axis in range(3):
            for side, val in ((0, mn[axis] - 1e-6), (1, mx[axis] + 1e-6)):
                pts = rng.uniform(mn, mx, size=(500, 3))
                pts[:, axis] = val
                assert not oracle_contains(s, pts).any(), (name, axis, side)


def test_bounds_follow_reference_formulas():
    cat = solid_catalogue()
    assert oracle_bounds(cat["sphere"]) == ((0.1 - 0.7, -0.2 - 0.7, 0.05 - 0.7), (0.1 + 0.7, -0.2 + 0.7, 0.05 + 0.7))
    # IntersectedSolid.Max clamps to Min (solid.go:159-160)
    disjoint = m3.IntersectedSolid([m3.Sphere((-2, 0, 0), 0.5), m3.Sphere((2, 0, 0), 0.5)])
    mn, mx = oracle_bounds(disjoint)
    assert mx[0] == mn[0] == 1.5
    # Scale with a negative factor swaps min and max (transform.go:75-81)
    flipped = m3.ScaleSolid(m3.Rect((0, 0, 0), (1, 2, 3)), -2.0)
    assert oracle_bounds(flipped) == ((-2.0, -4.0, -6.0), (0.0, 0.0, 0.0))


# ---------------------------------------------------------------------------
# Random solid manifoldness (TestMarchingCubesRandom, mc_test.go:27-42)
# ---------------------------------------------------------------------------
class HashSolid(m3.Solid):
    """A pure stand-in for the reference's randomSolid: inside the unit cube,
    a point is inside with probability 1/4 by a hash of its coordinate bits."""

    def __init__(self, seed):
        self.seed = seed

    def Min(self):
        return (0.0, 0.0, 0.0)

    def Max(self):
        return (1.0, 1.0, 1.0)

    def ContainsBatch(self, pts):
        pts = np.asarray(pts, dtype=np.float64)
        inb = np.all((pts >= 0) & (pts <= 1), axis=1)
        bits = np.ascontiguousarray(pts).view(np.uint64).astype(object)
        out = np.zeros(len(pts), dtype=bool)
        for i in range(len(pts)):
            h = (int(bits[i, 0]) * 0x9E3779B185EBCA87 ^ int(bits[i, 1]) * 0xC2B2AE3D27D4EB4F ^ int(bits[i, 2]) * 0x165667B19E3779F9 ^ self.seed) & ((1 << 64) - 1)
            h ^= h >> 29
            h = (h * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
            h ^= h >> 32
            out[i] = inb[i] and (h & 3) == 0
        return out


@pytest.mark.parametrize("iters", [0, 2])
def test_random_solid_closed_manifold(iters):
    for seed in range(6):
        tris, _ = oracle_search(HashSolid(seed), 0.1, iters, threads=4)
        assert len(tris) > 0
        directed_once, paired = edge_topology(tris)
        assert directed_once and paired, seed
        assert volume(tris) > 0  # outward orientation


# ---------------------------------------------------------------------------
# Errors (bounder.go:48-52, mc.go:687-689)
# ---------------------------------------------------------------------------
def test_invalid_bounds():
    with pytest.raises(OracleError) as e:
        oracle_search(m3.Sphere((0, 0, 0), float("nan")), 0.1, 0)
    assert e.value.code == 1


def test_inside_boundary():
    s = m3.ForceSolidBounds(m3.Sphere((0, 0, 0), 1.0), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    bad = m3.JoinedSolid([m3.Sphere((0, 0, 0), 0.2), _Unbounded()])
    with pytest.raises(OracleError) as e:
        oracle_search(bad, 0.1, 0)
    assert e.value.code == 2
    # clipped by its own bounds: fine
    tris, _ = oracle_search(s, 0.1, 0)
    assert len(tris) > 0


class _Unbounded(m3.Solid):
    """Reports small bounds but is true everywhere (violates the Solid contract)."""

    def Min(self):
        return (-0.3, -0.3, -0.3)

    def Max(self):
        return (0.3, 0.3, 0.3)

    def ContainsBatch(self, pts):
        return np.ones(len(pts), dtype=bool)


# ---------------------------------------------------------------------------
# ColliderSolid (collisions.go:119-134, primitives.go:207-249)
# ---------------------------------------------------------------------------
RAY_DIR = np.array([0.5224892708603626, 0.10494477243214506, 0.43558938446126527])


def two_boxes():
    """The two-box mesh of TestMarchingCubesFilter (mc_test.go:55-58)."""
    return m3.NewMeshRect((-1, -1, -1), (0, 0, 0)).AddMesh(m3.NewMeshRect((0.1, 0, 0), (1, 1, 1)))


def numpy_ray_parity(tris: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """Independent numpy restatement of ColliderContains: per point, count
    triangles hit by Triangle.rayCollision with scale >= 0 along RAY_DIR."""
    t = tris.reshape(-1, 3, 3)
    t0, t1, t2 = t[:, 0], t[:, 1], t[:, 2]
    n = np.cross(t1 - t0, t2 - t0)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    dn = RAY_DIR / np.linalg.norm(RAY_DIR)
    keep = ~((n @ dn < 1e-8) & (n @ dn > -1e-8))
    v1, v2 = t1 - t0, t2 - t0
    c1 = np.cross(RAY_DIR, v2)
    det = np.einsum("ij,ij->i", c1, v1)
    keep &= det != 0
    inv = 1 / np.where(det == 0, 1, det)
    out = np.zeros(len(pts), dtype=bool)
    for i, p in enumerate(pts):
        o = p - t0
        b1 = inv * np.einsum("ij,ij->i", o, c1)
        c2 = np.cross(o, v1)
        b2 = inv * (c2 @ RAY_DIR)
        sc = inv * np.einsum("ij,ij->i", v2, c2)
        hit = keep & ~((b1 < 0) | (b1 > 1)) & ~((b2 < 0) | (b1 + b2 > 1)) & ~(sc < 0)
        out[i] = hit.sum() % 2 == 1
    return out


def test_collider_two_boxes_closed_form():
    col = m3.NewColliderSolid(m3.MeshToCollider(two_boxes()))
    rng = np.random.default_rng(2)
    p = rng.uniform(-1.2, 1.2, size=(4000, 3))
    in1 = np.all((p > -1) & (p < 0), axis=1)
    in2 = np.all((p > [0.1, 0, 0]) & (p < 1), axis=1)
    d = np.minimum(np.abs(p - np.round(p * 10) / 10).min(axis=1), 1)  # distance to the faces' planes
    far = d > 1e-6
    got = oracle_contains(col, p)
    assert np.array_equal(got[far], (in1 | in2)[far])
    mn, mx = oracle_bounds(col)
    assert mn == (-1.0, -1.0, -1.0) and mx == (1.0, 1.0, 1.0)


def test_collider_matches_numpy_restatement():
    sph, _ = oracle_search(m3.Sphere((0.05, -0.02, 0.1), 0.8), 0.15, 4, threads=4)
    col = m3.NewColliderSolid(m3.MeshCollider(sph))
    rng = np.random.default_rng(4)
    p = rng.uniform(-1, 1, size=(600, 3))
    assert np.array_equal(oracle_contains(col, p), numpy_ray_parity(sph, p))
    bx = two_boxes().triangles().reshape(-1, 9)
    p = rng.uniform(-1.1, 1.1, size=(600, 3))
    assert np.array_equal(oracle_contains(m3.NewColliderSolid(m3.MeshCollider(bx)), p), numpy_ray_parity(bx, p))


# ---------------------------------------------------------------------------
# ColliderSolid with an inset or a hollow radius (solid.go:260-300,
# collisions.go:119-134, primitives.go:251-279)
# ---------------------------------------------------------------------------
def numpy_triangle_distance(tris: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Euclidean distance from p to each triangle (T, 9): the plane distance
    where the projection falls inside the triangle, else the nearest edge
    (the textbook decomposition, not the reference's ray formulation)."""
    t = tris.reshape(-1, 3, 3)
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    n = np.cross(b - a, c - a)
    nn = np.einsum("ij,ij->i", n, n)
    h = np.einsum("ij,ij->i", p - a, n) / nn
    q = p - h[:, None] * n
    inside = np.ones(len(t), dtype=bool)
    for u, v in ((a, b), (b, c), (c, a)):
        inside &= np.einsum("ij,ij->i", np.cross(v - u, q - u), n) >= 0
    d = np.where(inside, np.abs(h) * np.sqrt(nn), np.inf)
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        s = np.clip(np.einsum("ij,ij->i", p - u, e) / np.einsum("ij,ij->i", e, e), 0, 1)
        d = np.minimum(d, np.linalg.norm(u + s[:, None] * e - p, axis=1))
    return d


def test_triangle_sphere_collision_is_distance_below_radius():
    """Triangle.SphereCollision == (distance to the triangle < r) away from
    ties, on random triangles, centres and radii (including centres whose
    nearest feature is a vertex, an edge or the face)."""
    from oracle.binding import oracle_tri_sphere

    rng = np.random.default_rng(21)
    checked = {"face": 0, "other": 0}
    for _ in range(3000):
        tri = rng.normal(size=9)
        c = rng.normal(size=3) * rng.choice([0.3, 1.0, 3.0])
        d = numpy_triangle_distance(tri[None], c)[0]
        r = d * rng.choice([0.5, 0.99, 1.01, 2.0])
        if abs(d - r) < 1e-9 * (1 + r):
            continue
        assert oracle_tri_sphere(tri, c, r) == (d < r), (tri, c, r, d)
        t = tri.reshape(3, 3)
        n = np.cross(t[1] - t[0], t[2] - t[0])
        plane = abs((c - t[0]) @ n) / np.linalg.norm(n)
        checked["face" if abs(plane - d) < 1e-12 else "other"] += 1
    assert min(checked.values()) > 500


def _unit_box_distance(p):
    """(inside, distance to the surface of [0, 1]^3)."""
    inside = np.all((p > 0) & (p < 1), axis=1)
    d_in = np.minimum(p, 1 - p).min(axis=1)
    d_out = np.linalg.norm(np.maximum(0, np.maximum(-p, p - 1)), axis=1)
    return inside, np.where(inside, d_in, d_out)


def test_collider_inset_outset_hollow_on_a_box():
    """Closed forms on the unit box mesh: inset m keeps the points at least m
    inside; outset -m adds the points closer than m outside; the hollow shell
    of radius r is the points closer than r to the surface."""
    box = m3.NewMeshRect((0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(22)
    p = rng.uniform(-0.3, 1.3, size=(20000, 3))
    inside, d = _unit_box_distance(p)
    plane = np.abs(p - np.round(p)).min(axis=1)  # grazing rays along face planes
    m = 0.15
    cases = {
        "inset": (m3.NewColliderSolidInset(m3.MeshToCollider(box), m), inside & (d >= m)),
        "outset": (m3.NewColliderSolidInset(m3.MeshToCollider(box), -m), inside | (d < m)),
        "hollow": (m3.NewColliderSolidHollow(m3.MeshToCollider(box), m), d < m),
    }
    far = (np.abs(d - m) > 1e-9) & (plane > 1e-9)
    for name, (s, want) in cases.items():
        got = oracle_contains(s, p)
        assert np.array_equal(got[far], want[far]), (name, np.count_nonzero(got[far] != want[far]))
        assert want[far].sum() > 1000, name
    # bounds (solid.go:263-280): the inset box, max clamped to min when the
    # inset passes the middle, and the box grown by the radius
    assert oracle_bounds(cases["inset"][0]) == ((0.15,) * 3, (0.85,) * 3)
    assert oracle_bounds(m3.NewColliderSolidInset(m3.MeshToCollider(box), 0.7)) == ((0.7,) * 3, (0.7,) * 3)
    assert oracle_bounds(cases["outset"][0]) == ((-0.15,) * 3, (1.15,) * 3)
    assert oracle_bounds(cases["hollow"][0]) == ((-0.15,) * 3, (1.15,) * 3)


def test_collider_inset_on_a_sphere_mesh_matches_numpy():
    """Ray parity and the mesh distance, both restated in numpy, on a meshed
    sphere: inset keeps odd parity at distance >= m, outset adds even parity
    at distance < m, hollow is distance < r."""
    sph, _ = oracle_search(m3.Sphere((0.05, -0.02, 0.1), 0.8), 0.2, 4, threads=4)
    rng = np.random.default_rng(23)
    p = rng.uniform(-1, 1, size=(400, 3))
    parity = numpy_ray_parity(sph, p)
    d = np.array([numpy_triangle_distance(sph, q).min() for q in p])
    m = 0.1
    far = np.abs(d - m) > 1e-9
    col = m3.MeshCollider(sph)
    for s, want in ((m3.NewColliderSolidInset(col, m), parity & (d >= m)),
                    (m3.NewColliderSolidInset(col, -m), parity | (d < m)),
                    (m3.NewColliderSolidHollow(col, m), d < m)):
        got = oracle_contains(s, p)
        assert np.array_equal(got[far], want[far])
        assert 20 < want.sum() < len(p) - 20


# ---------------------------------------------------------------------------
# GroupTriangles (bvh.go:118-156, 219-292), the order of SaveGroupedSTL
# (mesh.go:721-743)
# ---------------------------------------------------------------------------
def test_group_triangles_split_properties():
    from oracle.binding import oracle_group_triangles

    rng = np.random.default_rng(31)
    for n in (0, 1, 2, 3, 5, 17, 300):
        tris = rng.normal(size=(n, 9))
        check_grouping(tris, oracle_group_triangles(tris))
    # a marching-cubes mesh: many equal midpoints (ties broken by index)
    mesh, _ = oracle_search(m3.Sphere((0.0, 0.0, 0.0), 0.6), 0.2, 0, threads=4)
    check_grouping(mesh, oracle_group_triangles(mesh))
    # a plausible slip -- the halves swapped, or the last best axis -- breaks it
    tris = rng.normal(size=(64, 9))
    perm = oracle_group_triangles(tris)
    with pytest.raises(AssertionError):
        check_grouping(tris, np.concatenate([perm[32:], perm[:32]]))


def test_new_mesh_rect_is_closed_and_outward():
    m = m3.NewMeshRect((0, 0, 0), (1, 2, 3))
    t = m.triangles()
    assert len(t) == 12 and m.NumTriangles() == 12
    directed_once, paired = edge_topology(t)
    assert directed_once and paired
    assert abs(volume(t) - 6.0) < 1e-12 and m.Volume() == 6.0


def test_csg64_oracle_smoke():
    s = workloads.csg64_solid(2)
    tris, info = oracle_search(s, 0.04, 8, threads=8)
    assert info["corners"] > 0 and len(tris) > 1000
    directed_once, paired = edge_topology(tris)
    assert directed_once and paired


# ---------------------------------------------------------------------------
# MarchingCubesInterior (mc.go:55-60, 191-198, 252-255)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("iters", [0, 1, 5])
def test_interior_points_on_sphere(iters):
    """Per vertex: the interior point is inside the unit sphere (textbook
    |p| <= 1, Go's Dist operation order), lies on the vertex's grid edge and
    is the bracket end half a final bracket away; the searched vertices and
    the triangles equal those of MarchingCubesSearch."""
    delta = 0.1
    s = m3.Sphere((0, 0, 0), 1.0)
    tris, info = oracle_search(s, delta, iters, threads=8, interior=True)
    plain, _ = oracle_search(s, delta, iters, threads=8)
    assert np.array_equal(canonical(tris), canonical(plain))
    pairs = info["interior"]
    assert len(pairs) == info["vertices"] == len(np.unique(tris.reshape(-1, 3), axis=0))
    assert {tuple(r) for r in pairs[:, :3].tolist()} == {tuple(r) for r in tris.reshape(-1, 3).tolist()}
    ax = oracle_spacer(s, delta)
    step = ax[0][1] - ax[0][0]
    for v, ip in zip(pairs[:, :3], pairs[:, 3:]):
        x, y, z = ip
        assert math.sqrt(x * x + y * y + z * z) <= 1.0
        diff = [i for i in range(3) if v[i] != ip[i]]
        assert len(diff) == 1  # the same grid edge, moved along its axis only
        a = diff[0]
        assert 0 < abs(v[a] - ip[a]) <= step / 2 ** (iters + 1) * (1 + 1e-9)


def test_interior_points_on_csg_are_inside():
    s = solid_catalogue()["csg"]
    _, info = oracle_search(s, 0.05, 4, threads=8, interior=True)
    pts = info["interior"][:, 3:]
    assert oracle_contains(s, pts).all()
    # the searched vertex itself is not guaranteed inside; the interior point is
    assert len(pts) > 100


# ---------------------------------------------------------------------------
# MarchingCubesConj (mc.go:66-71): the inverse map of the vertices
# ---------------------------------------------------------------------------
def test_inverse_apply_elementary_ops():
    """Inverse().Apply per op (transform.go:52-54, 83-85, 114-116, 153-155):
    translate subtracts, scale divides, a rotation's inverse is its
    transpose, and the joined inverse runs the ops backwards."""
    rng = np.random.default_rng(9)
    p = rng.uniform(-2, 2, size=(50, 3))
    t = np.array([0.25, -0.5, 0.125])
    assert np.array_equal(oracle_inverse_apply(m3.Translate(t), p), p + (-t))
    assert np.array_equal(oracle_inverse_apply(m3.Scale(4.0), p), p * 0.25)
    assert np.array_equal(oracle_inverse_apply(m3.VecScale((2.0, 4.0, 0.5)), p), p * np.array([0.5, 0.25, 2.0]))
    rot = m3.Rotation((0.0, 0.6, 0.8), 0.7)
    R = np.array(rot.Matrix).reshape(3, 3)
    assert np.allclose(oracle_inverse_apply(rot, p), p @ R, atol=1e-12)  # R^T p per row
    # joined: Translate then Scale; the inverse scales back first, then subtracts
    j = m3.JoinedTransform([m3.Translate(t), m3.Scale(2.0)])
    assert np.array_equal(oracle_inverse_apply(j, p), p * 0.5 + (-t))


def test_conj_power_of_two_scale_is_exact():
    """Conj by Scale(2) meshes the doubled sphere and halves its vertices
    exactly: x2 gives back the plain search of the doubled sphere, bit for
    bit (a forward instead of inverse map would be 4x off)."""
    s = m3.Sphere((0.1, -0.2, 0.05), 0.6)
    got, _ = oracle_conj(s, 0.05, 3, m3.Scale(2.0), threads=8)
    big, _ = oracle_search(m3.TransformSolid(m3.Scale(2.0), s), 0.05, 3, threads=8)
    assert np.array_equal(canonical(got * 2.0), canonical(big))


def test_conj_rotation_stays_on_surface():
    """Conj by a rotation and translation: every vertex within one grid step
    of the original unit sphere (closed form)."""
    s = m3.Sphere((0, 0, 0), 1.0)
    delta = 0.08
    got, _ = oracle_conj(s, delta, 4, m3.Rotation((1.0, 0.0, 0.0), 0.4), m3.Translate((0.3, 0.0, -0.2)), threads=8)
    r = np.linalg.norm(got.reshape(-1, 3), axis=1)
    assert np.all(np.abs(r - 1.0) <= delta)
    assert abs(volume(got) - 4 * math.pi / 3) < 0.05


# ---------------------------------------------------------------------------
# EncodeSTL (export.go:20-66, fileformats/stl.go:17-58)
# ---------------------------------------------------------------------------
def _parse_stl(data):
    """Binary STL reader (stdlib struct): header, count, (normals, vertices, trailers)."""
    n = struct.unpack_from("<I", data, 80)[0]
    assert len(data) == 84 + 50 * n
    recs = [struct.unpack_from("<12fH", data, 84 + 50 * i) for i in range(n)]
    arr = np.array([r[:12] for r in recs], dtype=np.float32).reshape(n, 4, 3)
    return data[:80], n, arr[:, 0], arr[:, 1:], [r[12] for r in recs]


def test_stl_known_answer():
    tri = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0]])
    want = bytes(80) + struct.pack("<I", 1) + struct.pack("<12f", 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0) + bytes(2)
    assert oracle_encode_stl(tri) == want
    assert oracle_encode_stl(np.zeros((0, 9))) == bytes(80) + struct.pack("<I", 0)


def test_stl_round_trip_and_normals():
    """The reference's TestImportSTL / TestSTLBinary: the vertices read back
    are the float32 casts; normals are unit, orthogonal to the edges and
    right-handed (t1 - t0) x (t2 - t0); a degenerate triangle's normal is NaN
    (1 / 0 scaled by 0)."""
    rng = np.random.default_rng(3)
    tris = rng.normal(size=(200, 9))
    tris[7] = np.tile(tris[7][:3], 3)  # degenerate
    header, n, normals, verts, trailers = _parse_stl(oracle_encode_stl(tris))
    assert header == bytes(80) and n == 200 and all(t == 0 for t in trailers)
    assert np.array_equal(verts.reshape(-1, 9), tris.astype(np.float32))
    assert np.isnan(normals[7]).all()
    ok = np.arange(200) != 7
    t = tris[ok].reshape(-1, 3, 3)
    e1, e2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    cr = np.cross(e1, e2)
    nn = normals[ok].astype(np.float64)
    assert np.allclose(np.linalg.norm(nn, axis=1), 1, atol=1e-6)
    scale = np.linalg.norm(e1, axis=1) + np.linalg.norm(e2, axis=1)
    assert (np.abs(np.sum(nn * e1, axis=1)) <= 1e-5 * scale).all()
    assert (np.abs(np.sum(nn * e2, axis=1)) <= 1e-5 * scale).all()
    assert (np.sum(nn * cr, axis=1) > 0).all()


# ---------------------------------------------------------------------------
# MarchingCubesFilter / SearchFilter / C2F (mc.go:73-180)
# ---------------------------------------------------------------------------
def test_rect_collision_closed_form():
    """Triangle.RectCollision over a cube's surface (NewMeshRect): a rect
    crossing a face collides, one strictly inside or outside does not, one
    enclosing the cube does (the triangles' edges lie in it)."""
    f = oracle_rect_collider(m3.NewMeshRect((0, 0, 0), (1, 1, 1)).triangles())
    assert f((0.5, 0.5, -0.1), (0.6, 0.6, 0.1))
    assert f((0.9, 0.2, 0.2), (1.1, 0.3, 0.3))
    assert not f((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))
    assert not f((1.1, 1.1, 1.1), (2, 2, 2))
    assert f((-1, -1, -1), (2, 2, 2))
    # a rect edge piercing a face, no triangle edge inside the rect
    assert f((0.45, 0.45, -0.5), (0.55, 0.55, 0.5))


def test_library_rect_collider_matches_oracle():
    """The library's rect collider (a bounding-volume tree, host code) and the
    oracle's brute force agree on random rects around a sphere mesh."""
    tris, _ = oracle_search(m3.Sphere((0.1, 0, 0), 0.8), 0.1, 4, threads=8)
    lib_c = m3.MeshToCollider(tris)
    brute = oracle_rect_collider(tris)
    rng = np.random.default_rng(4)
    hits = 0
    for _ in range(3000):
        c = rng.uniform(-1.2, 1.2, size=3)
        h = rng.uniform(0.0, 0.3, size=3)
        mn, mx = tuple(c - h), tuple(c + h)
        want = brute(mn, mx)
        hits += want
        assert lib_c.RectCollision(m3.Rect(mn, mx)) == want
    assert 300 < hits < 2700


def test_filter_always_true_is_marching_cubes():
    s = solid_catalogue()["csg"]
    for it in (0, 3):
        base, _ = oracle_search(s, 0.05, it, threads=8)
        got = oracle_search_filter(s, lambda mn, mx: True, 0.05, it)
        assert np.array_equal(canonical(got), canonical(base))
    assert len(oracle_search_filter(s, lambda mn, mx: False, 0.05, 3)) == 0


def test_filter_collider_solid_equals_mc():
    """TestMarchingCubesFilter (mc_test.go:44-67): a collider solid filtered
    by its own collider's RectCollision equals plain MarchingCubes."""
    sphere_mesh, _ = oracle_search(m3.Sphere((0.1, 0.3, -0.2), 1.0), 0.15, 3, threads=8)
    for mesh in (sphere_mesh, np.concatenate([m3.NewMeshRect((-1, -1, -1), (0, 0, 0)).triangles().reshape(-1, 9),
                                               m3.NewMeshRect((0.1, 0, 0), (1, 1, 1)).triangles().reshape(-1, 9)])):
        solid = m3.NewColliderSolid(m3.MeshToCollider(mesh))
        base, _ = oracle_search(solid, 0.1, 0, threads=8)
        got = oracle_search_filter(solid, oracle_rect_collider(mesh), 0.1, 0)
        assert np.array_equal(canonical(got), canonical(base))


def test_filter_half_space_keeps_only_its_cubes():
    """A filter accepting blocks with min x < 0: every triangle lies in a cube
    whose low x corner is below the largest accepted block edge, and the
    triangles are a subset of the unfiltered ones."""
    s = m3.Sphere((0, 0, 0), 1.0)
    base, _ = oracle_search(s, 0.1, 0, threads=8)
    got = oracle_search_filter(s, lambda mn, mx: mn[0] < 0, 0.1, 0)
    assert 0 < len(got) < len(base)
    bset = {tuple(r) for r in canonical(base).tolist()}
    assert all(tuple(r) in bset for r in canonical(got).tolist())
    assert got.reshape(-1, 3)[:, 0].min() < -0.9


def test_c2f_equals_search():
    """TestMarchingCubesC2F (mc_test.go:69-94)."""
    s = m3.Sphere((0.1, 0.3, -0.2), 1.0)
    a, _ = oracle_search(s, 0.05, 8, threads=8)
    assert np.array_equal(canonical(oracle_c2f(s, 0.1, 0.05, 0, 8)), canonical(a))
    mesh = np.concatenate([m3.NewMeshRect((-1, -1, -1), (0, 0, 0)).triangles().reshape(-1, 9),
                           m3.NewMeshRect((0.3, 0, 0), (1, 1, 1)).triangles().reshape(-1, 9)])
    boxes = m3.NewColliderSolid(m3.MeshToCollider(mesh))
    a, _ = oracle_search(boxes, 0.1, 8, threads=8)
    b = oracle_c2f(boxes, 0.2, 0.1, 0, 8)
    assert np.array_equal(canonical(b), canonical(a))
    ok, _ = edge_topology(b)
    assert ok
