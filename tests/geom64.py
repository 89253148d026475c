"""Independent float64 geometry used to pin the oracle (textbook formulations,
not restatements of the reference): Sutherland-Hodgman polygon clipping,
barycentric point-in-triangle, slab-interval box test, Moller-Trumbore."""
from __future__ import annotations

import numpy as np


def clip_polygon_box(poly, lo, hi):
    """Sutherland-Hodgman clip of a convex polygon (k,3) against an axis-aligned box."""
    pts = [np.asarray(p, np.float64) for p in poly]
    for axis in range(3):
        for side in (0, 1):
            if not pts:
                return []
            bound = lo[axis] if side == 0 else hi[axis]
            inside = (lambda p: p[axis] >= bound) if side == 0 else (lambda p: p[axis] <= bound)
            out = []
            for i in range(len(pts)):
                cur, prev = pts[i], pts[i - 1]
                ci, pi = inside(cur), inside(prev)
                if ci:
                    if not pi:
                        t = (bound - prev[axis]) / (cur[axis] - prev[axis])
                        out.append(prev + t * (cur - prev))
                    out.append(cur)
                elif pi:
                    t = (bound - prev[axis]) / (cur[axis] - prev[axis])
                    out.append(prev + t * (cur - prev))
            pts = out
    return pts


def polygon_area(pts):
    if len(pts) < 3:
        return 0.0
    p0 = pts[0]
    s = np.zeros(3)
    for i in range(1, len(pts) - 1):
        s += np.cross(pts[i] - p0, pts[i + 1] - p0)
    return 0.5 * float(np.linalg.norm(s))


def polygon_centroid(pts):
    p0 = pts[0]
    acc = np.zeros(3)
    wsum = 0.0
    for i in range(1, len(pts) - 1):
        a = 0.5 * np.linalg.norm(np.cross(pts[i] - p0, pts[i + 1] - p0))
        acc += a * (p0 + pts[i] + pts[i + 1]) / 3.0
        wsum += a
    return acc / wsum if wsum > 0 else np.mean(pts, axis=0)


def tri_area(t9):
    t = np.asarray(t9, np.float64).reshape(3, 3)
    return 0.5 * float(np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])))


def barycentric(tri9, p):
    t = np.asarray(tri9, np.float64).reshape(3, 3)
    a, b, c = t
    v0, v1, v2 = b - a, c - a, np.asarray(p, np.float64) - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    return np.array([1.0 - v - w, v, w])


def moller_trumbore(tri9, o, d):
    """Returns t (float64) of the ray/triangle-plane hit inside the triangle, or None."""
    t = np.asarray(tri9, np.float64).reshape(3, 3)
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    e1, e2 = t[1] - t[0], t[2] - t[0]
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-300:
        return None
    inv = 1.0 / det
    s = o - t[0]
    u = (s @ p) * inv
    q = np.cross(s, e1)
    v = (d @ q) * inv
    tt = (e2 @ q) * inv
    if u < 0 or v < 0 or u + v > 1 or tt <= 0:
        return None
    return tt


def slab_interval(lo, hi, o, d):
    """(tmin, tmax) of the line o + t d inside the box, float64, or None if parallel-outside."""
    tmin, tmax = -np.inf, np.inf
    for k in range(3):
        if d[k] == 0.0:
            if o[k] < lo[k] or o[k] > hi[k]:
                return None
            continue
        t0 = (lo[k] - o[k]) / d[k]
        t1 = (hi[k] - o[k]) / d[k]
        tmin = max(tmin, min(t0, t1))
        tmax = min(tmax, max(t0, t1))
    return tmin, tmax
