"""The C-ABI library loads, exports every declared symbol, and its host-side
logic (Solid.Min/Max, case table) matches the oracle.  No GPU needed: none of
these calls launches a kernel."""
import json
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


def test_case_table_matches_mc_go_fixture():
    """The library's edge-id table against the table derived from mc.go's
    text (tests/golden/make_case_table_from_mc_go.py): each corner pair of the
    reference's mcTriangle is the cube edge the library names by axis and
    the two other corner bits."""
    with open(os.path.join(ROOT, "tests", "golden", "mc_case_table_mc_go.json")) as f:
        ref = [c["triangles"] for c in json.load(f)["cases"]]

    def edge_id(a, b):
        d, lo = a ^ b, a & b
        assert d in (1, 2, 4)
        axis = d.bit_length() - 1
        others = [bit for bit in range(3) if bit != axis]
        return 4 * axis + ((lo >> others[0]) & 1) + 2 * ((lo >> others[1]) & 1)

    lib_t = m3.case_table()
    for k in range(256):
        assert int(lib_t[k, 0]) == len(ref[k]), k
        want = [edge_id(t[2 * e], t[2 * e + 1]) for t in ref[k] for e in range(3)]
        assert [int(x) for x in lib_t[k, 1 : 1 + 3 * len(ref[k])]] == want, k
    # corner 0 alone: edges x, y, z at the origin corner, in that order
    assert [int(x) for x in lib_t[1, :4]] == [1, 0, 4, 8]


def test_plan_fold_skip_equals_serial_fold():
    """The brick plan's subtree-skipping fold against the instruction-by-
    instruction fold on 20,000 random programs x 16 random bricks (host
    code, no GPU): identical folded code, root and overflow flag."""
    import ctypes

    L = _lib.lib()
    checked, bad = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.mc3d_selftest_plan_fold(20240607, 20000, ctypes.byref(checked), ctypes.byref(bad)) == 0
    assert checked.value == 20000 * 16
    assert bad.value == 0
