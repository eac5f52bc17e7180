"""Restatement of the reference's hierarchical fast winding number in 2D --
TEST INFRASTRUCTURE ONLY.

The library decides containment with the exact winding sum (the oracle's
orc_winding).  The reference's titwcsph builds its containment function with
geom::MakeFastWinding (titwcsph/wcsph.cpp:76-77): a tree over the faces whose
far nodes are replaced by a third-order expansion, with contains() repeating a
query near the threshold at twice the accuracy and finally with the exact sum
(geom/winding/fast_winding.hpp:80-92).  This file restates that estimator
(fast_winding.hpp:35-374) so that tests can check, point by point, that the
exact sum classifies every point as the reference's estimator does.

Pinned by the reference's own test (geom/winding.test.cpp:125-146: estimates
within 6e-3 of the exact sum on a tessellated unit square, both orientations,
and the same classification), and by the exact sum for leaf-only trees.
"""
from __future__ import annotations

import math

import numpy as np

_TWO_PI = 2.0 * math.pi  # unit_sphere_area_v<2>


def segment_winding(a, b, p) -> float:
    """Segment::winding_number (geom/segment.hpp:76-80)."""
    ap = (a[0] - p[0], a[1] - p[1])
    bp = (b[0] - p[0], b[1] - p[1])
    det = ap[0] * bp[1] - ap[1] * bp[0]
    return math.atan2(det, ap[0] * bp[0] + ap[1] * bp[1]) / _TWO_PI


def exact_winding(verts, faces, p) -> float:
    """ExactWindingFunc::operator() (geom/winding/exact_winding.hpp): the sum
    over every face."""
    return sum(segment_winding(verts[a], verts[b], p) for a, b in faces)


class _Agg:
    """Aggregate_ (fast_winding.hpp:96-104): box, weight, weighted centre and
    the three moments about the origin."""

    __slots__ = ("lo", "hi", "weight", "wc", "m0", "m1", "m2")

    @staticmethod
    def face(a, b) -> "_Agg":
        """make_aggregate_ (:190-225) of one segment: wnormal = cross(b - a)
        = (ba_y, -ba_x), centre (a + b) / 2, second moment
        wn_k ((a_i + b_i)(a_j + b_j) + a_i a_j + b_i b_j) / 6."""
        g = _Agg()
        a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
        g.lo, g.hi = np.minimum(a, b), np.maximum(a, b)
        ba = b - a
        wn = np.array([ba[1], -ba[0]])
        c = 0.5 * (a + b)
        g.weight = math.hypot(wn[0], wn[1])
        g.wc = g.weight * c
        g.m0 = wn.copy()
        g.m1 = np.outer(c, wn)                                   # [i][j] = wn_j c_i
        s = np.outer(a + b, a + b) + np.outer(a, a) + np.outer(b, b)
        g.m2 = s[:, :, None] * wn[None, None, :] / 6.0          # Dim (Dim + 1) = 6
        return g

    def join(self, o: "_Agg") -> "_Agg":
        """combine_aggregates_ (:228-259)."""
        g = _Agg()
        g.lo, g.hi = np.minimum(self.lo, o.lo), np.maximum(self.hi, o.hi)
        g.weight = self.weight + o.weight
        g.wc = self.wc + o.wc
        g.m0, g.m1, g.m2 = self.m0 + o.m0, self.m1 + o.m1, self.m2 + o.m2
        return g


class _Node:
    __slots__ = ("center", "radius_sqr", "m0", "m1", "m2", "left", "right", "first", "last")


class FastWinding:
    """FastWindingFunc<Vec<double, 2>> with MakeFastWinding's defaults
    (leaf size 8, beta 2; fast_winding.hpp:383-386)."""

    def __init__(self, verts, faces, leaf_size: int = 8, beta: float = 2.0):
        self.verts = np.asarray(verts, dtype=np.float64)
        self.faces = np.asarray(faces, dtype=np.int64).reshape(-1, 2)
        self.leaf_size, self.beta = int(leaf_size), float(beta)
        self.perm = np.arange(len(self.faces))
        self.nodes: list[_Node] = []
        if len(self.faces):
            self._build(0, len(self.faces))

    def _center(self, f):
        a, b = self.faces[f]
        return 0.5 * (self.verts[a] + self.verts[b])

    def _build(self, lo: int, hi: int):
        """build_node_ (:124-176): a leaf at most leaf_size faces, else a median
        split of the face centres along the axis of their box's largest extent
        (nth_element: the lower half holds the smallest centres)."""
        k = len(self.nodes)
        self.nodes.append(_Node())
        idx = self.perm[lo:hi]
        if hi - lo <= self.leaf_size:
            agg = None
            for f in idx:
                a, b = self.faces[f]
                g = _Agg.face(self.verts[a], self.verts[b])
                agg = g if agg is None else agg.join(g)
            node = self.nodes[k]
            node.first, node.last, node.left, node.right = lo, hi, 0, 0
        else:
            cs = np.array([self._center(f) for f in idx])
            ext = cs.max(0) - cs.min(0)
            axis = int(np.argmax(ext))  # max_value_index: the first largest
            mid = (hi - lo) // 2
            order = np.argsort(cs[:, axis], kind="stable")
            self.perm[lo:hi] = idx[order]
            left, la = self._build(lo, lo + mid)
            right, ra = self._build(lo + mid, hi)
            agg = la.join(ra)
            node = self.nodes[k]
            node.first = node.last = 0
            node.left, node.right = left, right
        self._init(self.nodes[k], agg)
        return k, agg

    @staticmethod
    def _init(node: _Node, agg: _Agg):
        """init_node_ (:262-291): moments about the weighted centre."""
        c = agg.wc / agg.weight if agg.weight > 0.0 else 0.5 * (agg.lo + agg.hi)
        node.center = c
        r = np.maximum(np.abs(c - agg.lo), np.abs(agg.hi - c))
        node.radius_sqr = float(r @ r)
        node.m0 = agg.m0.copy()
        node.m1 = agg.m1 - np.outer(c, agg.m0)
        node.m2 = (agg.m2 - c[:, None, None] * agg.m1[None, :, :]
                   - c[None, :, None] * agg.m1[:, None, :]
                   + np.outer(c, c)[:, :, None] * agg.m0[None, None, :])

    @staticmethod
    def _expansion(d, d2, m0, m1, m2) -> float:
        """eval_expansion_ (:326-368), Dim = 2: derivatives of the 2D
        Green's function gradient d / (2 pi |d|^2) to third order."""
        scale = 1.0 / _TWO_PI
        inv2 = 1.0 / d2
        inv_dim = math.sqrt(inv2) ** 2
        r = 0.0
        for i in range(2):
            r += m0[i] * (scale * inv_dim * d[i])
        for i in range(2):
            for j in range(2):
                ident = 1.0 if i == j else 0.0
                r += m1[i][j] * (scale * inv_dim * (ident - 2.0 * d[i] * d[j] * inv2))
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    idt = ((d[k] if i == j else 0.0) + (d[j] if i == k else 0.0)
                           + (d[i] if j == k else 0.0))
                    third = 2.0 * scale * inv_dim * inv2 * (4.0 * d[i] * d[j] * d[k] * inv2 - idt)
                    r += 0.5 * m2[i][j][k] * third
        return r

    def estimate(self, p, beta: float | None = None) -> float:
        """eval_fast_ (:301-323)."""
        if not self.nodes:
            return 0.0
        b2 = (self.beta if beta is None else beta) ** 2
        p = np.asarray(p, dtype=np.float64)
        out, stack = 0.0, [0]
        while stack:
            node = self.nodes[stack.pop()]
            d = node.center - p
            d2 = float(d @ d)
            if d2 > b2 * node.radius_sqr:
                out += self._expansion(d, d2, node.m0, node.m1, node.m2)
            elif node.first != node.last:
                for f in self.perm[node.first:node.last]:
                    a, b = self.faces[f]
                    out += segment_winding(self.verts[a], self.verts[b], p)
            else:
                stack.append(node.right)
                stack.append(node.left)
        return out

    def contains(self, p, threshold: float = 0.5, uncertainty: float = 1e-3):
        """contains (:80-92): the estimate, again at 2 beta when it lies within
        `uncertainty` of the threshold, then the exact sum.  Returns the
        classification and how many refinements it took (0, 1 or 2)."""
        w = self.estimate(p)
        tries = 0
        if abs(w - threshold) < uncertainty:
            w, tries = self.estimate(p, 2.0 * self.beta), 1
            if abs(w - threshold) < uncertainty:
                w, tries = exact_winding(self.verts, self.faces, p), 2
        return w > threshold, tries
