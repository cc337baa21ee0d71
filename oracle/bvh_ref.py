"""TEST INFRASTRUCTURE ONLY: pure-Python restatement of the reference's BVH
builders in float32 scalar arithmetic, for small inputs.

build_blas follows src/scene/blas.rs:44-314 and build_tlas src/scene/tlas.rs
:29-263 (recursive subdivide, 20 / 2 SAH bins, leaf at <= 8 / <= 1 items,
bins and running boxes starting from the origin-anchored AABB::default()).
It pins the C++ builder of libpulse_hip (pulse_build_blas / pulse_build_tlas),
which must emit the same node arrays byte for byte.
"""
from __future__ import annotations

import numpy as np

F = np.float32
FMAX = F(np.finfo(np.float32).max)


def _vmin(a, b):
    return [F(min(x, y)) for x, y in zip(a, b)]


def _vmax(a, b):
    return [F(max(x, y)) for x, y in zip(a, b)]


class _Aabb:
    def __init__(self):
        self.mn = [F(0), F(0), F(0)]
        self.mx = [F(0), F(0), F(0)]

    def grow(self, p):
        self.mn = _vmin(self.mn, p)
        self.mx = _vmax(self.mx, p)

    def grow_aabb(self, b):
        self.grow(b.mn)
        self.grow(b.mx)

    def area(self):
        e = [F(self.mx[i] - self.mn[i]) for i in range(3)]
        return F(F(F(e[0] * e[1]) + F(e[1] * e[2])) + F(e[2] * e[0]))


def _as_usize(f):
    f = float(f)
    if not (f > 0):
        return 0
    return int(f)


class _Builder:
    def __init__(self, centroids, item_points, leaf_max, bins):
        self.c = centroids          # list of [F, F, F]
        self.pts = item_points      # list of lists of points grown into bins / node boxes
        self.leaf_max = leaf_max
        self.bins = bins
        self.nodes = []             # [mn, first, mx, count]
        self.idx = list(range(len(centroids)))

    def node_aabb(self, first, count):
        mn = [FMAX] * 3
        mx = [-FMAX] * 3
        for i in range(count):
            for p in self.pts[self.idx[first + i]]:
                mn = _vmin(mn, p)
                mx = _vmax(mx, p)
        return mn, mx

    def split(self, first, count):
        best = (0, F(0.0), F(1e30))
        BC = self.bins
        for axis in range(3):
            bmin, bmax = F(1e30), F(-1e30)
            for i in range(count):
                v = self.c[self.idx[first + i]][axis]
                bmin = F(min(bmin, v))
                bmax = F(max(bmax, v))
            if bmin == bmax:
                continue
            boxes = [_Aabb() for _ in range(BC)]
            counts = [0] * BC
            inv = F(F(BC) / F(bmax - bmin))
            for i in range(count):
                it = self.idx[first + i]
                b = min(BC - 1, _as_usize(F(F(self.c[it][axis] - bmin) * inv)))
                counts[b] += 1
                for p in self.pts[it]:
                    boxes[b].grow(p)
            area_a, area_b = [F(0)] * (BC - 1), [F(0)] * (BC - 1)
            cnt_a, cnt_b = [0] * (BC - 1), [0] * (BC - 1)
            ba, bb = _Aabb(), _Aabb()
            sa = sb = 0
            for i in range(BC - 1):
                sa += counts[i]
                cnt_a[i] = sa
                ba.grow_aabb(boxes[i])
                area_a[i] = ba.area()
                sb += counts[BC - 1 - i]
                cnt_b[BC - 2 - i] = sb
                bb.grow_aabb(boxes[BC - 1 - i])
                area_b[BC - 2 - i] = bb.area()
            size = F(F(bmax - bmin) / F(BC))
            for i in range(BC - 1):
                cost = F(F(F(cnt_a[i]) * area_a[i]) + F(F(cnt_b[i]) * area_b[i]))
                if cost < best[2]:
                    best = (axis, F(bmin + F(size * F(i + 1))), cost)
        return best

    def subdivide(self, ni):
        first, count = self.nodes[ni][1], self.nodes[ni][3]
        if count <= self.leaf_max:
            return
        axis, pos, _ = self.split(first, count)
        i, j = first, first + count - 1
        while i <= j:
            if self.c[self.idx[i]][axis] < pos:
                i += 1
            else:
                self.idx[i], self.idx[j] = self.idx[j], self.idx[i]
                j -= 1
        a_count = i - first
        if a_count == 0 or a_count == count:
            return
        ca = len(self.nodes)
        mn, mx = self.node_aabb(first, a_count)
        self.nodes.append([mn, first, mx, a_count])
        mn, mx = self.node_aabb(i, count - a_count)
        self.nodes.append([mn, i, mx, count - a_count])
        self.nodes[ni][1] = ca
        self.nodes[ni][3] = 0
        self.subdivide(ca)
        self.subdivide(ca + 1)

    def run(self):
        n = len(self.c)
        mn, mx = self.node_aabb(0, n)
        self.nodes.append([mn, 0, mx, n])
        self.subdivide(0)
        return self.nodes, self.idx


def _pack(nodes, dtype, first_name, count_name):
    out = np.zeros(len(nodes), dtype)
    for k, (mn, first, mx, count) in enumerate(nodes):
        out[k]["aabb_min"] = mn
        out[k]["aabb_max"] = mx
        out[k][first_name] = first
        out[k][count_name] = count
    return out


def build_blas(prims: np.ndarray, blas_dtype) -> tuple:
    """prims: structured array with p_first/p_second/p_third (f32x3)."""
    pts, cents = [], []
    for p in prims:
        tri = [[F(v) for v in p["p_first"]], [F(v) for v in p["p_second"]], [F(v) for v in p["p_third"]]]
        mn, mx = [FMAX] * 3, [-FMAX] * 3
        for q in tri:
            mn = _vmin(mn, q)
        for q in tri:
            mx = _vmax(mx, q)
        cents.append([F(mn[i] + F(F(0.5) * F(mx[i] - mn[i]))) for i in range(3)])
        pts.append(tri)
    nodes, idx = _Builder(cents, pts, 8, 20).run()
    return _pack(nodes, blas_dtype, "a_or_first_tri", "tri_count"), np.array(idx, np.uint32)


def build_tlas(bounds_min, bounds_max, centers, tlas_dtype) -> tuple:
    n = len(centers)
    if n == 0:
        return np.zeros(1, tlas_dtype), np.zeros(0, np.uint32)
    pts = [[[F(v) for v in bounds_min[i]], [F(v) for v in bounds_max[i]]] for i in range(n)]
    cents = [[F(v) for v in centers[i]] for i in range(n)]
    nodes, idx = _Builder(cents, pts, 1, 2).run()
    return _pack(nodes, tlas_dtype, "a_or_first_instance", "instance_count"), np.array(idx, np.uint32)
