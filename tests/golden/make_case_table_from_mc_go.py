"""Writes mc_case_table_mc_go.json: the 256-case marching cubes table derived
from the text of the reference's model3d/mc.go, independently of the two
restatements it checks (the library's build_case_table and the oracle's).

The 23 base cases are parsed from `baseTriangleTable` (mc.go:462-586) as
data; the rotation group and the lookup rule are restated here in Python
from mc.go:312-353 (allMcRotations: closure of the two generators by a
queue, then a lexicographic sort), mc.go:356-388 (Compose, ApplyCorner,
ApplyTriangle, ApplyIntersections) and mc.go:434-456 (mcLookupTable: for each
base case, the first rotation in sorted order that produces a key decides
its triangles).  Each triangle is kept as the reference stores it: six
corner indices, two per cube edge, in vertex order.

Run where the reference checkout exists (the build container only; the
fixture is committed and the tests read the fixture):

    python tests/golden/make_case_table_from_mc_go.py /root/reference/model3d/mc.go
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_base_table(text):
    """{intersections bits: [6-corner triangles]} from baseTriangleTable."""
    start = text.index("var baseTriangleTable = map[mcIntersections][]mcTriangle{")
    body = text[start:]
    end = body.index("\n}\n")
    body = body[:end]
    table = {}
    key = None
    for line in body.split("\n")[1:]:
        line = line.split("//")[0].strip()
        if not line:
            continue
        m = re.match(r"newMcIntersections\(([\d,\s]*)\):\s*\{(\},)?$", line)
        if m:
            corners = [int(c) for c in m.group(1).replace(" ", "").split(",") if c != ""]
            key = 0
            for c in corners:
                key |= 1 << c
            assert key not in table, key
            table[key] = []
            if m.group(2):  # `{},` on the same line: no triangles
                key = None
            continue
        m = re.match(r"\{(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\},$", line)
        if m:
            assert key is not None
            table[key].append([int(g) for g in m.groups()])
            continue
        if line == "},":
            key = None
            continue
        raise ValueError(f"unparsed line in baseTriangleTable: {line!r}")
    return table


def compose(m, m1):
    """mcRotation.Compose: res[i] = m[m1[i]] (mc.go:356-362)."""
    return tuple(m[m1[i]] for i in range(8))


def all_rotations():
    """allMcRotations (mc.go:312-353)."""
    z_rot = (2, 0, 3, 1, 6, 4, 7, 5)
    x_rot = (2, 3, 6, 7, 0, 1, 4, 5)
    queue = [tuple(range(8))]
    seen = {queue[0]}
    while queue:
        nxt = queue.pop(0)
        for op in (z_rot, x_rot):
            r = compose(op, nxt)
            if r not in seen:
                seen.add(r)
                queue.append(r)
    return sorted(seen)  # lexicographic, as the sort.Slice comparator


def apply_intersections(rot, bits):
    """mcRotation.ApplyIntersections (mc.go:377-388)."""
    res = 0
    for c in range(8):
        if bits & (1 << c):
            res |= 1 << rot[c]
    return res


def lookup_table(base):
    """mcLookupTable (mc.go:434-456).  Every key is reached from one base
    case only (checked), so the order of the base cases does not matter."""
    rots = all_rotations()
    result, owner = {}, {}
    for bits, tris in base.items():
        for rot in rots:
            key = apply_intersections(rot, bits)
            if key in owner and owner[key] != bits:
                raise AssertionError(f"key {key} reached from base cases {owner[key]} and {bits}")
            owner[key] = bits
            if key not in result:
                result[key] = [[rot[c] for c in t] for t in tris]
    assert sorted(result) == list(range(256))
    return result, len(rots)


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/model3d/mc.go"
    with open(path) as f:
        text = f.read()
    base = parse_base_table(text)
    assert len(base) == 23, len(base)
    table, nrot = lookup_table(base)
    assert nrot == 24
    out = {
        "source": "derived from the text of model3d/mc.go: baseTriangleTable (462-586) parsed as data, "
        "allMcRotations / Compose / ApplyIntersections / mcLookupTable (312-388, 434-456) restated in Python",
        "base_cases": len(base),
        "rotations": nrot,
        "cases": [{"case": k, "triangles": table[k]} for k in range(256)],
    }
    with open(os.path.join(HERE, "mc_case_table_mc_go.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {sum(len(c['triangles']) for c in out['cases'])} triangles over 256 cases")


if __name__ == "__main__":
    main()
