"""Containment: the exact winding sum that the library uses against the
reference's fast winding estimator (geom/winding/fast_winding.hpp), CPU.

titwcsph decides containment with MakeFastWinding over its pool polygon
(titwcsph/wcsph.cpp:66-77); the library and the oracle use the exact sum of
Segment::winding_number instead.  oracle/fast_winding.py restates the
estimator; these tests pin that restatement with the reference's own winding
test and then check, near the threshold 1/2, that the exact sum classifies
every point as the estimator's contains() does: on the titwcsph pool, on
tessellated polygons whose trees use the far-field expansion, at distances
from 1e-12 to 1e-1 of a face on both sides, on faces and at corners.
"""
import math

import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from oracle.fast_winding import FastWinding, exact_winding


def _square(reverse=False):
    """make_square of geom/winding.test.cpp:12-30."""
    v = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    f = [(3, 0), (2, 3), (1, 2), (0, 1)] if reverse else [(0, 1), (1, 2), (2, 3), (3, 0)]
    return v, np.array(f)


def _l_shape(d_max):
    v = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 0.7], [0.8, 0.7], [0.8, 1.5], [0.0, 1.5]])
    f = np.array([(k, (k + 1) % 6) for k in range(6)])
    return ta.tessellate(v, f, d_max)


@pytest.mark.parametrize("reverse", [False, True])
def test_restatement_passes_the_reference_winding_test(reverse):
    """geom/winding.test.cpp:125-146: the unit square tessellated to 0.05,
    estimates within 6e-3 of the exact sum, same classification."""
    v, f = ta.tessellate(*_square(reverse), 0.05)
    fw = FastWinding(v, f)
    assert len(fw.nodes) > 1  # a real tree: 80 faces, leaves of 8
    for p in [(0.1, 0.1), (0.5, 0.5), (0.9, 0.75), (-0.5, 0.5), (0.5, 1.5), (2.0, 2.0),
              (100.0, 100.0)]:
        exact = exact_winding(v, f, p)
        assert abs(fw.estimate(p) - exact) <= 6e-3
        assert fw.contains(p)[0] == (exact > 0.5)
        # the oracle's exact sum is the same sum
        assert abs(oracle.winding(v, f, [p])[0] - exact) <= 1e-15


def test_restatement_expansion_converges():
    """Far from an open chain of faces (half the L-shape) the third-order
    expansion's error falls as (size / distance)^4: a mistake in a moment or
    a derivative would leave a lower power.  (A closed polygon's far field
    vanishes term by term, which would pin nothing.)"""
    v, f = _l_shape(0.05)
    f = f[: len(f) // 2]
    fw = FastWinding(v, f, leaf_size=1000)  # one node: the expansion alone
    errs = []
    for dist in (8.0, 16.0, 32.0):
        p = (1.0 + dist * math.cos(0.3), 0.75 + dist * math.sin(0.3))
        node = fw.nodes[0]
        d = node.center - np.asarray(p)
        errs.append(abs(fw._expansion(d, float(d @ d), node.m0, node.m1, node.m2)
                        - exact_winding(v, f, p)))
    assert errs[0] / errs[1] > 12.0 and errs[1] / errs[2] > 12.0, errs


def _near_points(v, f, rng, n_random=4000):
    """Points at distances 1e-12 .. 1e-1 (times the face length) on both sides of
    each face's midpoint and quarter points, on the faces, at and around the
    corners, at the vertices, and random points over the grown bounding box."""
    pts = []
    for a, b in f:
        A, B = v[a], v[b]
        t = B - A
        L = math.hypot(*t)
        n = np.array([t[1], -t[0]]) / L
        for s in (0.25, 0.5, 0.75):
            q = A + s * t
            pts.append(q)
            for e in range(1, 13):
                for sgn in (1.0, -1.0):
                    pts.append(q + sgn * 10.0 ** (-e) * L * n)
        pts.append(A.copy())  # a tessellation vertex on a straight side: exactly 1/2
        for e in range(2, 13):
            for ang in np.linspace(0.0, 2.0 * math.pi, 8, endpoint=False):
                pts.append(A + 10.0 ** (-e) * np.array([math.cos(ang), math.sin(ang)]))
    lo, hi = v.min(0), v.max(0)
    span = hi - lo
    pts.extend(rng.uniform(lo - 0.5 * span, hi + 0.5 * span, size=(n_random, 2)))
    return np.asarray(pts)


def _classify(v, f, pts):
    fw = FastWinding(v, f)
    exact = oracle.winding(v, f, pts)
    fast, tries = [], np.zeros(3, dtype=int)
    for p in pts:
        c, k = fw.contains(p)
        fast.append(c)
        tries[k] += 1
    return np.asarray(fast), exact, tries, fw


def test_pool_classification_is_identical_everywhere():
    """titwcsph's containment polygon (wcsph.cpp:66-75, four faces: one leaf,
    so every query within twice its radius is the exact sum in face order):
    identical classification at every point, the wall vertices (where the
    fixed particles sit, on the polygon) included."""
    walls = ta.box_walls(3.22, 1.8, 0.02)
    v, f = np.asarray(walls.cverts), np.asarray(walls.cfaces, dtype=np.int64)
    rng = np.random.default_rng(7)
    pts = np.concatenate([_near_points(v, f, rng), np.asarray(walls.verts)])
    fast, exact, tries, _ = _classify(v, f, pts)
    print(f"pool: {len(pts)} points, refinements {tries.tolist()}")
    assert np.array_equal(fast, exact > 0.5)


@pytest.mark.parametrize("case", ["square", "l_shape"])
def test_tessellated_classification_differs_only_on_the_boundary(case):
    """Tessellated polygons (trees with far-field nodes): identical at every
    point off the polygon, down to 1e-12 of a face's length.  Points that lie
    on it at a vertex inside a straight side have an exact sum of 1/2 (up to
    rounding; outside when it rounds to 1/2 or below), while the estimate there carries the
    far nodes' error (about 1.6e-3 on the tessellated unit square), outside
    the 1e-3 band that triggers refinement, so the estimator may call them
    inside.  Such points are the only differences."""
    v, f = ta.tessellate(*_square(), 0.05) if case == "square" else _l_shape(0.05)
    f = np.asarray(f, dtype=np.int64)
    rng = np.random.default_rng(7)
    pts = _near_points(v, f, rng, n_random=1500)
    fast, exact, tries, fw = _classify(v, f, pts)
    differ = fast != (exact > 0.5)
    on_boundary = np.abs(exact - 0.5) < 1e-12  # 1/2 up to the sum's rounding
    worst = max(abs(fw.estimate(p) - w) for p, w in zip(pts, exact) if abs(w - 0.5) > 0.25)
    print(f"{case}: {len(pts)} points, refinements {tries.tolist()}, {differ.sum()} differ "
          f"(all on the boundary: {bool(np.all(on_boundary[differ]))}), "
          f"largest estimate error off the boundary {worst:.1e}")
    assert np.all(on_boundary[differ])
    assert tries[2] > 0     # some on-boundary points do reach the exact sum
    assert worst < 0.05


def test_a_coarse_estimator_would_misclassify():
    """With the acceptance ratio cut to 0.3 and no refinement the estimator
    misclassifies points near the L-shape: the comparison above can tell."""
    v, f = _l_shape(0.05)
    fw = FastWinding(v, f, beta=0.3)
    rng = np.random.default_rng(3)
    pts = _near_points(v, np.asarray(f, dtype=np.int64), rng, n_random=500)
    exact = oracle.winding(v, f, pts)
    bad = sum(fw.contains(p, uncertainty=0.0)[0] != (w > 0.5) for p, w in zip(pts, exact))
    assert bad > 0
