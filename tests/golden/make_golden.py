"""Writes the committed fixtures of tests/golden/.

  mc_case_table.json   the 256-case table (mcLookupTable, model3d/mc.go:434-456)
                       as produced by the oracle; pinned by the properties in
                       tests/test_oracle.py (determinism, single base case per
                       key, sign-changing edges, closed oriented output).
  spacer_sphere_r1_d0.02.json
                       Xs of newSquareSpacer (mc.go:594-608) for
                       Sphere{Radius: 1} at delta 0.02, computed here with
                       plain Python floats (IEEE binary64 adds, as in Go).
  sphere_r1_d0.02_counts.json
                       closed-form counts of the sphere workload.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def spacer(lo, hi, delta):
    out = []
    x = lo - delta
    while x <= hi + delta:
        out.append(x)
        x += delta
    return out


def main():
    from oracle.binding import oracle_case_table

    table, conflicts = oracle_case_table()
    assert conflicts == 0
    rows = []
    for k in range(256):
        n = int(table[k, 0])
        rows.append({"case": k, "triangles": [[int(c) for c in table[k, 1 + 6 * t : 7 + 6 * t]] for t in range(n)]})
    with open(os.path.join(HERE, "mc_case_table.json"), "w") as f:
        json.dump({"source": "oracle mcLookupTable restatement", "cases": rows}, f, indent=0)

    xs = spacer(-1.0, 1.0, 0.02)
    with open(os.path.join(HERE, "spacer_sphere_r1_d0.02.json"), "w") as f:
        json.dump(
            {
                "lo": -1.0,
                "hi": 1.0,
                "delta": 0.02,
                "values_hex": [struct.pack(">d", v).hex() for v in xs],
                "count": len(xs),
                "last": repr(xs[-1]),
            },
            f,
            indent=0,
        )

    with open(os.path.join(HERE, "sphere_r1_d0.02_counts.json"), "w") as f:
        json.dump({"corners": 102**3, "vertices": 46998, "triangles": 93992, "bisection_evals_iters8": 46998 * 9}, f)


if __name__ == "__main__":
    main()
