"""Regenerates tests/golden/search_points.json: the inputs of the reference's
search test (gen_search_points.cpp) and, as expected outputs, the neighbour
sets of an O(N^2) brute force in exact rational arithmetic (fractions), so
that no floating-point convention of this build enters the expected values.
The reference's naive search uses a strict '<' (search.test.cpp:48); both the
strict and the inclusive sets are stored (they coincide for these points).
"""
import json
import subprocess
import sys
import tempfile
from fractions import Fraction
from pathlib import Path

HERE = Path(__file__).resolve().parent
RADII = [0.01, 0.1, 0.5, 1.0]


def brute(points, radius, strict):
    r2 = Fraction(radius) ** 2
    P = [[Fraction(c) for c in p] for p in points]
    out = []
    for a, pa in enumerate(P):
        row = []
        for b, pb in enumerate(P):
            s = sum((x - y) ** 2 for x, y in zip(pa, pb))
            if (s < r2) if strict else (s <= r2):
                row.append(b)
        out.append(row)
    return out


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "gen"
        subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe),
                        str(HERE / "gen_search_points.cpp")], check=True)
        data = json.loads(subprocess.run([str(exe)], check=True,
                                         capture_output=True, text=True).stdout)
    data["radii"] = RADII
    for key in ("points2d", "points3d"):
        pts = data[key]
        data[key + "_strict"] = {str(r): brute(pts, r, True) for r in RADII}
        incl = {str(r): brute(pts, r, False) for r in RADII}
        assert incl == data[key + "_strict"], "ties in the random points"
    (HERE / "search_points.json").write_text(json.dumps(data))
    print("wrote", HERE / "search_points.json", file=sys.stderr)


if __name__ == "__main__":
    main()
