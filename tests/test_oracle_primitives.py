"""Pins the oracle's geometric primitives against closed forms and independent
float64 geometry (tests/geom64.py), and against the reference's own numeric
invariant (Triangle.test.h:13-42).  A transposed operand, a wrong sign or a
dropped clipping case in the oracle fails at least one test here."""
import numpy as np
import pytest

import geom64 as G

F = np.float32


def T9(a, b, c):
    return np.array(list(a) + list(b) + list(c), np.float32)


# --------------------------------------------------------------- Triangle::cast
def test_tri_cast_closed_forms(oracle):
    tri = T9((0, 0, 5), (2, 0, 5), (0, 2, 5))           # n = (b-a)x(c-a) = (0,0,4)
    hit, t, p = oracle.tri_cast(tri, (0.5, 0.5, 0), (0, 0, 1), clip=True)
    assert hit and t == F(5.0) and tuple(p) == (0.5, 0.5, 5.0)       # front face: origin on the -n side
    # from the +n side: clipped as a backface, hit without clipping (Triangle.impl.h:337)
    assert not oracle.tri_cast(tri, (0.5, 0.5, 10), (0, 0, -1), clip=True)[0]
    hit, t, _ = oracle.tri_cast(tri, (0.5, 0.5, 10), (0, 0, -1), clip=False)
    assert hit and t == F(5.0)
    # t is parametric along an unnormalised direction
    hit, t, _ = oracle.tri_cast(tri, (0.5, 0.5, 0), (0, 0, 2), clip=True)
    assert hit and t == F(2.5)
    # exactly through a vertex: containsVertex early-out (Triangle.impl.h:293)
    hit, t, p = oracle.tri_cast(tri, (2, 0, 0), (0, 0, 1), clip=True)
    assert hit and tuple(p) == (2.0, 0.0, 5.0)
    # parallel, behind, outside
    assert not oracle.tri_cast(tri, (0.5, 0.5, 0), (1, 0, 0))[0]
    assert not oracle.tri_cast(tri, (0.5, 0.5, 0), (0, 0, -1))[0]
    assert not oracle.tri_cast(tri, (1.5, 1.5, 0), (0, 0, 1))[0]
    # degenerate triangles never hit (Triangle.impl.h:292)
    for deg in (T9((0, 0, 5), (0, 0, 5), (0, 2, 5)), T9((0, 0, 5), (2, 0, 5), (2, 0, 5)),
                T9((0, 0, 5), (2, 0, 5), (0, 0, 5))):
        assert not oracle.tri_cast(deg, (0.0, 0.0, 0), (0, 0, 1))[0]


def test_tri_cast_vs_moller_trumbore(oracle):
    rng = np.random.default_rng(1)
    agree = 0
    for _ in range(6000):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = rng.uniform(-1, 1, 3)
        d = (target - o).astype(np.float32)
        t64 = G.moller_trumbore(tri, o, d)
        hit, t, p = oracle.tri_cast(tri, o, d, clip=False)
        bary = None
        if t64 is not None:
            bary = G.barycentric(tri, o.astype(np.float64) + t64 * d.astype(np.float64))
        if t64 is not None and bary.min() > 1e-3 and G.tri_area(tri) > 1e-3:
            assert hit, (tri, o, d)
            assert abs(float(t) - t64) <= 1e-4 * max(1.0, abs(t64))
            agree += 1
        elif t64 is None:
            # clearly outside (barycentric margin) -> the oracle must miss too
            t_plane = None
            n = np.cross(tri[3:6].astype(np.float64) - tri[0:3], tri[6:9].astype(np.float64) - tri[0:3])
            den = n @ d
            if abs(den) > 1e-9:
                t_plane = (n @ (tri[0:3].astype(np.float64) - o)) / den
            if t_plane is not None and t_plane > 1e-3:
                b = G.barycentric(tri, o + t_plane * d.astype(np.float64))
                if b.min() < -1e-2:
                    assert not hit
    assert agree > 300


def test_backface_sign_convention(oracle):
    rng = np.random.default_rng(2)
    for _ in range(500):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        a, b, c = tri[0:3].astype(np.float64), tri[3:6].astype(np.float64), tri[6:9].astype(np.float64)
        n = np.cross(b - a, c - a)
        if np.linalg.norm(n) < 1e-2:
            continue
        cen = (a + b + c) / 3
        side = 2.0 * n / np.linalg.norm(n)
        for sgn, front in ((-1, True), (1, False)):     # origin on the -n side sees the front face
            o = (cen + sgn * side).astype(np.float32)
            d = (cen - o).astype(np.float32)
            assert oracle.tri_cast(tri, o, d, clip=True)[0] == front
            assert oracle.tri_cast(tri, o, d, clip=False)[0]


def _angles(tri):
    a, b, c = tri[0:3].astype(np.float64), tri[3:6].astype(np.float64), tri[6:9].astype(np.float64)
    ang = lambda u, v: float(np.arccos(np.clip(u @ v / (np.linalg.norm(u) * np.linalg.norm(v)), -1, 1)))
    return a, b, c, ang


def test_contains_vs_barycentric(oracle):
    """Points clearly inside are contained and points clearly outside are not, for
    triangles whose smallest angle is at least 10 degrees (so the 32*eps tolerance,
    which acts on cosines, stays far below the 1e-2 barycentric margin)."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(4000):
        tri = rng.uniform(-2, 2, 9).astype(np.float32)
        if G.tri_area(tri) < 0.05:
            continue
        a, b, c, ang = _angles(tri)
        if min(ang(b - a, c - a), ang(c - b, a - b), ang(a - c, b - c)) < np.radians(10):
            continue
        w = rng.uniform(-0.3, 1.3, 3)
        w /= w.sum()
        p = (w[0] * tri[0:3] + w[1] * tri[3:6] + w[2] * tri[6:9]).astype(np.float32)
        bc = G.barycentric(tri, p)
        if bc.min() > 1e-2:
            assert oracle.tri_contains(tri, p)
            checked += 1
        elif bc.min() < -1e-2:
            assert not oracle.tri_contains(tri, p)
            checked += 1
    assert checked > 1000


def test_contains_is_the_angle_cone_test(oracle):
    """Triangle.impl.h:299-301 accepts x iff at each vertex v with outgoing edge e
    (ab at a, bc at b, ca at c) the angle between e and x - v is at most
    arccos(cos(interior angle at v) - 32*eps/|e|).  Checked with float64 arccos angles,
    including thin triangles where the tolerance is large."""
    rng = np.random.default_rng(13)
    tol = float(np.float32(32) * np.float32(5e-6))
    checked = near_edge = 0
    for _ in range(8000):
        tri = rng.uniform(-2, 2, 9).astype(np.float32)
        if G.tri_area(tri) < 1e-3:
            continue
        a, b, c, ang = _angles(tri)
        w = rng.uniform(-0.2, 1.2, 3)
        w /= w.sum()
        p = (w[0] * tri[0:3] + w[1] * tri[3:6] + w[2] * tri[6:9]).astype(np.float32)
        x = p.astype(np.float64)
        margins = []
        for v, e, f in ((a, b - a, c - a), (b, c - b, a - b), (c, a - c, b - c)):
            lim = np.arccos(np.clip(np.cos(ang(e, f)) - tol / np.linalg.norm(e), -1, 1))
            margins.append(lim - ang(e, x - v))
        got = oracle.tri_contains(tri, p)
        if min(margins) > 1e-4:
            assert got
            checked += 1
            near_edge += int(G.barycentric(tri, p).min() < -1e-3)
        elif min(margins) < -1e-4:
            assert not got
            checked += 1
    assert checked > 3000 and near_edge > 0


# --------------------------------------------------------------- AABB
def test_slab_vs_interval(oracle):
    rng = np.random.default_rng(4)
    n_hit = n_miss = 0
    for _ in range(5000):
        lo = rng.uniform(-2, 1, 3)
        hi = lo + rng.uniform(0.1, 2, 3)
        box = np.concatenate([lo, hi]).astype(np.float32)
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        if rng.random() < 0.5:                      # aim at a point inside the box
            d = (rng.uniform(lo, hi) - o).astype(np.float32)
        inv = (np.float32(1.0) / d).astype(np.float32)
        r = oracle.aabb_cast_preinv(box, o, inv)
        iv = G.slab_interval(box[:3].astype(np.float64), box[3:].astype(np.float64), o.astype(np.float64),
                             d.astype(np.float64))
        tmin, tmax = iv
        if tmax - tmin > 1e-3:
            assert r != np.finfo(np.float32).max
            assert abs(float(r) - tmin) <= 1e-5 * max(1.0, abs(tmin))
            n_hit += 1
        elif tmin - tmax > 1e-3:
            assert r == np.finfo(np.float32).max
            n_miss += 1
    assert n_hit > 500 and n_miss > 500


def test_slab_has_no_t_cull(oracle):
    """Boxes behind the origin still count as hit (AABB.impl.h:128-129, clip=false)."""
    box = np.array([-1, -1, -10, 1, 1, -5], np.float32)
    o = np.zeros(3, np.float32)
    d = np.array([1e-3, 1e-3, 1.0], np.float32)
    r = oracle.aabb_cast_preinv(box, o, (np.float32(1.0) / d).astype(np.float32))
    assert r < 0 and r != np.finfo(np.float32).max


def _rand_box_tri(rng):
    lo = rng.uniform(-1, 0.5, 3)
    hi = lo + rng.uniform(0.2, 1.5, 3)
    c = rng.uniform(-1.2, 1.5, 3)
    tri = (c + rng.uniform(-0.8, 0.8, (3, 3))).astype(np.float32)
    return np.concatenate([lo, hi]).astype(np.float32), tri.reshape(9)


def test_intersects_tri_vs_polygon_clip(oracle):
    rng = np.random.default_rng(5)
    yes = no = 0
    for _ in range(4000):
        box, tri = _rand_box_tri(rng)
        poly = tri.astype(np.float64).reshape(3, 3)
        lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
        inner = G.clip_polygon_box(poly, lo + 1e-3, hi - 1e-3)
        outer = G.clip_polygon_box(poly, lo - 1e-3, hi + 1e-3)
        got = oracle.aabb_intersects_tri(box, tri)
        if len(inner) >= 3:
            assert got, (box, tri)
            yes += 1
        elif len(outer) == 0:
            assert not got, (box, tri)
            no += 1
    assert yes > 300 and no > 300


def test_clip_pieces_area_bounds(oracle):
    rng = np.random.default_rng(6)
    nonempty = 0
    for _ in range(3000):
        box, tri = _rand_box_tri(rng)
        pieces, n = oracle.clip_tri(box, tri)
        assert n <= 27
        lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
        poly = G.clip_polygon_box(tri.astype(np.float64).reshape(3, 3), lo, hi)
        area = G.polygon_area(poly) if len(poly) >= 3 else 0.0
        got = sum(G.tri_area(p) for p in pieces)
        assert abs(got - area) <= 1e-4 * max(area, 1e-3) + 1e-6, (box, tri, got, area)
        b = oracle.intersection_bounds(box, tri)
        if n > 0:
            v = pieces.reshape(-1, 3).astype(np.float64)
            assert np.all(v >= lo - 1e-5) and np.all(v <= hi + 1e-5)
            if area > 1e-4:
                P = np.array(poly)
                assert np.allclose(b[:3], P.min(0), atol=1e-5) and np.allclose(b[3:], P.max(0), atol=1e-5)
                nonempty += 1
        else:
            assert np.array_equal(b, box)          # empty clip returns the box itself (Shapes.impl.cuh:290)
    assert nonempty > 200


def test_intersection_center(oracle):
    rng = np.random.default_rng(7)
    for _ in range(2000):
        box, tri = _rand_box_tri(rng)
        c = oracle.intersection_center(box, tri)
        lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
        poly = G.clip_polygon_box(tri.astype(np.float64).reshape(3, 3), lo, hi)
        pieces, n = oracle.clip_tri(box, tri)
        if n == 0:
            assert np.array_equal(c, ((box[:3] + box[3:]) * np.float32(0.5)).astype(np.float32))
            continue
        area = G.polygon_area(poly) if len(poly) >= 3 else 0.0
        if area > 0.05:
            # each piece carries +1e-4 of weight (Shapes.impl.cuh:301), so the centre is the
            # area-weighted centroid up to ~n*1e-4/area of the box size
            tol = 2e-3 + 4 * n * 1e-4 / area
            assert np.allclose(c, G.polygon_centroid(poly), atol=tol * float(np.max(hi - lo)))


def test_get_masses_roundtrip(oracle):
    """Triangle.test.h:13-42: massCenter(getMasses(x)) within 2.5 of x for coordinates < 99999."""
    rng = np.random.default_rng(8)
    worst = 0.0
    for _ in range(20000):
        t = rng.integers(0, 99999, 9).astype(np.float32)
        a, b, c = t[0:3].astype(np.float64), t[3:6].astype(np.float64), t[6:9].astype(np.float64)
        n = np.cross(b - a, c - a)
        if np.linalg.norm(n) == 0:
            continue
        q = rng.integers(0, 99999, 3).astype(np.float64)
        x = (q + n * ((a - q) @ n) / (n @ n)).astype(np.float32)   # projectVertex (Triangle.impl.h:243-245)
        m = oracle.tri_get_masses(t, x)
        y = oracle.tri_mass_center(t, m)
        err = float(np.linalg.norm(y.astype(np.float64) - x))
        worst = max(worst, err)
    assert worst <= 2.5, worst


def test_constants():
    assert np.float32(256) * np.float32(5e-6) == np.float32(0.00128)
    assert np.float32(32) * np.float32(5e-6) == np.float32(0.00016)
