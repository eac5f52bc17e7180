"""Pins of the oracle's neighbour search (CPU).

The reference's own test (source/tit/geom/search.test.cpp:41-163) compares
the grid index against an O(N^2) search on 200 random points from
std::mt19937(123) for radii {0.01, 0.1, 0.5, 1} with cell hints 0.5R and 5R.
The points and the exact-arithmetic expected sets live in
tests/golden/search_points.json (made by tests/golden/make_search_fixture.py).
Lattices add exact ties at the radius, where the inclusive predicate of
BSphere::contains (bsphere.hpp:52-54) matters.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = json.loads((Path(__file__).parent / "golden" / "search_points.json").read_text())


def _lists(off, idx):
    return [list(idx[off[a]:off[a + 1]]) for a in range(off.size - 1)]


@pytest.mark.parametrize("key", ["points2d", "points3d"])
@pytest.mark.parametrize("radius", [0.01, 0.1, 0.5, 1.0])
@pytest.mark.parametrize("scale", [0.5, 5.0])
def test_grid_search_matches_reference_fixture(key, radius, scale):
    pts = np.array(GOLDEN[key])
    want = GOLDEN[key + "_strict"][str(radius)]
    off, idx = oracle.grid_search(pts, radius, scale * radius)
    got = _lists(off, idx)
    assert got == want  # ascending lists, self included


def test_reference_points_reproduced():
    # spot values of std::mt19937(123) + uniform_real_distribution(0, 1)
    assert GOLDEN["points2d"][0] == [0.7129553215802809, 0.42847092502893397]
    assert len(GOLDEN["points2d"]) == 200 and len(GOLDEN["points3d"]) == 200


def test_empty_input():
    off, idx = oracle.grid_search(np.zeros((0, 2)), 0.1, 0.05)
    assert off.tolist() == [0] and idx.size == 0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("hint_factor", [0.5, 1.0, 2.3])
def test_lattice_ties_inclusive(dim, hint_factor):
    dr = 0.01
    n = 18 if dim == 2 else 9
    g = np.meshgrid(*[np.arange(n)] * dim, indexing="ij")
    pts = np.stack([dr * (gi.ravel() + 1.0) for gi in g], axis=1)
    R = 4 * dr
    got = oracle.grid_search(pts, R, hint_factor * R)
    want = oracle.naive_search(pts, R)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # interior points see the whole lattice ball |(i, j[, k])| <= 4
    cnt = np.diff(got[0]).max()
    assert cnt == (49 if dim == 2 else 257)
    # Some pairs sit exactly on R^2 in FP64 (others miss it by an ulp), so the
    # strict and inclusive predicates give different sets: ties are exercised.
    strict = oracle.naive_search(pts, R, strict=True)
    assert strict[1].size < got[1].size


def test_naive_matches_numpy_bruteforce():
    rng = np.random.default_rng(0)
    pts = rng.uniform(size=(300, 3))
    R = 0.2
    off, idx = oracle.naive_search(pts, R)
    d = pts[None, :, :] - pts[:, None, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    for a in range(300):
        assert list(idx[off[a]:off[a + 1]]) == list(np.nonzero(s[a] <= R * R)[0])
