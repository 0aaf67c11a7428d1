"""The C-ABI library loads, exports every declared symbol, and its host-side
logic (Solid.Min/Max, case table) matches the oracle.  No GPU needed: none of
these calls launches a kernel."""
import os
import re

import numpy as np

import model3d_amd as m3
from model3d_amd import _lib, workloads
from oracle.binding import oracle_bounds, oracle_case_table
from tests.helpers import solid_catalogue

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "mc3d.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mc3d_[a-z0-9_]+)\s*\(", text)) - {"mc3d_contains_fn"})


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for name in syms:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name
    assert set(_lib.SIGNATURES) == set(syms)
    assert L.mc3d_version().startswith(b"mc3d")


def test_bounds_match_oracle_bitwise():
    cat = solid_catalogue()
    cat["csg64"] = workloads.csg64_solid(2)
    cat["neg_scale"] = m3.ScaleSolid(m3.Rect((0, 0, 0), (1, 2, 3)), -2.0)
    cat["disjoint_inter"] = m3.IntersectedSolid([m3.Sphere((-2, 0, 0), 0.5), m3.Sphere((2, 0, 0), 0.5)])
    for name, s in cat.items():
        got = (tuple(s.Min()), tuple(s.Max()))
        want = oracle_bounds(s)
        g = np.array(got).view(np.uint64)
        w = np.array(want).view(np.uint64)
        assert np.array_equal(g, w), name


def test_case_table_matches_oracle():
    lib_t = m3.case_table()
    or_t, _ = oracle_case_table()
    for k in range(256):
        n = int(or_t[k, 0])
        assert lib_t[k, 0] == n
        for t in range(n):
            for e in range(3):
                a, b = int(or_t[k, 1 + 6 * t + 2 * e]), int(or_t[k, 2 + 6 * t + 2 * e])
                d, lo = a ^ b, a & b
                if d == 1:
                    want = (lo >> 1) & 3
                elif d == 2:
                    want = 4 + ((lo & 1) | (((lo >> 2) & 1) << 1))
                else:
                    want = 8 + (lo & 3)
                assert lib_t[k, 1 + 3 * t + e] == want, (k, t, e)


def test_constructor_errors_are_reported():
    L = _lib.lib()
    assert not L.mc3d_joined(None, 0)
    assert "at least one" in _lib.last_error()
