"""Seeded synthetic solids for the benchmark and the parity tests.

Go's math/rand stream is not available outside Go, so the generators use
splitmix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): state += 0x9E3779B97F4A7C15, then the xor-shift
multiply finaliser; a float64 in [0, 1) is the top 53 bits times 2^-53.
"""
from __future__ import annotations

import math

import numpy as np

from .solids import (
    Capsule,
    ClipSolid,
    Cylinder,
    IntersectedSolid,
    JoinedSolid,
    Rect,
    Sphere,
    SubtractedSolid,
)

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)

    def unit(self):
        while True:
            v = [self.uniform(-1, 1) for _ in range(3)]
            n = math.sqrt(sum(x * x for x in v))
            if 1e-3 < n <= 1:
                return [x / n for x in v]


def random_primitive(rng: SplitMix64, box: float = 1.0):
    """One of the four primitive kinds with a centre uniform in [-box, box]^3."""
    c = [rng.uniform(-box, box) for _ in range(3)]
    kind = rng.randint(0, 3)
    if kind == 0:
        return Sphere(c, rng.uniform(0.05, 0.3))
    if kind == 1:
        side = [rng.uniform(0.1, 0.5) for _ in range(3)]
        return Rect([c[i] - side[i] / 2 for i in range(3)], [c[i] + side[i] / 2 for i in range(3)])
    d = rng.unit()
    half = rng.uniform(0.05, 0.3)
    p1 = [c[i] - d[i] * half for i in range(3)]
    p2 = [c[i] + d[i] * half for i in range(3)]
    r = rng.uniform(0.03, 0.2)
    return Capsule(p1, p2, r) if kind == 2 else Cylinder(p1, p2, r)


def _centre(s) -> list:
    if isinstance(s, Sphere):
        return list(s.Center)
    if isinstance(s, Rect):
        return [(s.MinVal[i] + s.MaxVal[i]) / 2 for i in range(3)]
    return [(s.P1[i] + s.P2[i]) / 2 for i in range(3)]


def random_csg(seed: int, n_prims: int = 64, p_union: float = 0.7, p_sub: float = 0.2, clip: float | None = 1.0):
    """A random binary CSG tree over ``n_prims`` primitives.

    Bottom-up: pick a random node, pair it with the node whose centre is
    nearest (so subtractions and intersections act on overlapping parts, as
    in a modelled part), and join the pair by union (70%), subtraction (20%)
    or intersection (10%).  With ``clip`` the root is ClipSolid'ed to
    [-clip, clip]^3."""
    rng = SplitMix64(seed)
    prims = [random_primitive(rng) for _ in range(n_prims)]
    nodes = [(p, _centre(p)) for p in prims]
    while len(nodes) > 1:
        i = rng.randint(0, len(nodes) - 1)
        a, ca = nodes.pop(i)
        # nearest remaining centre (first of equals): ((dx^2 + dy^2) + dz^2)
        d = np.array([c for _, c in nodes], dtype=np.float64) - np.array(ca, dtype=np.float64)
        d = d * d
        j = int(np.argmin((d[:, 0] + d[:, 1]) + d[:, 2]))
        b, cb = nodes.pop(j)
        u = rng.uniform()
        if u < p_union:
            n = JoinedSolid([a, b])
        elif u < p_union + p_sub:
            n = SubtractedSolid(a, b)
        else:
            n = IntersectedSolid([a, b])
        nodes.append((n, [(ca[t] + cb[t]) / 2 for t in range(3)]))
    root = nodes[0][0]
    if clip is not None:
        root = ClipSolid(root, (-clip,) * 3, (clip,) * 3)
    return root


# The benchmark configurations, by name.
SPHERE_CONFIG = dict(name="sphere_r1_d0.02", delta=0.02, iters=8)
CSG64_CONFIG = dict(name="csg64_seed2_d0.005", seed=2, n_prims=64, delta=0.005, iters=8)


COLLIDER_CONFIG = dict(name="collider_csg64_d0.02_256cubed", seed=2, n_prims=64, mesh_delta=0.02, iters=8)


# Many solids of uneven size in one call: primitive counts ~U{4..128}, grid
# edges ~U{32..256} cells, iters ~U{4..8}, each a clipped random CSG tree.
BATCH_CONFIG = dict(name="batch256_seed4", seed=4, n_solids=256, prims=(4, 128), cells=(32, 256), iters=(4, 8))


def batch_solids(seed: int = 4, n_solids: int = 256, prims=(4, 128), cells=(32, 256), iters=(4, 8)):
    """(solids, deltas, iters) of the batch workload: solid i is
    random_csg(its own seed, U{prims}) and its delta gives U{cells} cells
    along its longest bounds axis."""
    rng = SplitMix64(seed)
    solids, deltas, its = [], [], []
    for _ in range(n_solids):
        n_prims = rng.randint(*prims)
        n_cells = rng.randint(*cells)
        it = rng.randint(*iters)
        s = random_csg(rng.next_u64(), n_prims)
        mn, mx = s.Min(), s.Max()
        extent = max(mx[k] - mn[k] for k in range(3))
        solids.append(s)
        deltas.append(extent / n_cells)
        its.append(it)
    return solids, deltas, its


# One large solid: a 512-primitive tree in [-1, 1]^3 at delta 0.002
# (1002^3 ~ 1.0G corners, ~50M triangles).
LARGE_CONFIG = dict(name="csg512_seed9_d0.002", seed=9, n_prims=512, delta=0.002, iters=8)


def large_solid(seed: int = 9, n_prims: int = 512):
    return random_csg(seed, n_prims)


def sphere_solid():
    return Sphere((0.0, 0.0, 0.0), 1.0)


def collider_solid(seed: int = 2, n_prims: int = 64, mesh_delta: float = 0.02, samples: int = 256):
    """The CSG workload meshed at ``mesh_delta`` (on the GPU), wrapped as
    NewColliderSolid(MeshToCollider(mesh)); returns (solid, delta) with delta
    giving ``samples`` grid samples along the longest axis."""
    from .mc import MarchingCubesSearch
    from .solids import MeshCollider, NewColliderSolid

    mesh = MarchingCubesSearch(csg64_solid(seed, n_prims), mesh_delta, 8)
    solid = NewColliderSolid(MeshCollider(mesh.triangles()))
    mn, mx = solid.Min(), solid.Max()
    extent = max(mx[i] - mn[i] for i in range(3))
    return solid, extent / (samples - 3), len(mesh.faces)


def csg64_solid(seed: int = 2, n_prims: int = 64):
    return random_csg(seed, n_prims)
