// fp32 geometric primitives of the DumbRay hot path, shared by the host octree
// builder and the gfx950 kernels.  Every function evaluates the reference's
// expression tree in the same order with one correctly-rounded IEEE-754 fp32
// operation per source operation (all translation units are compiled with
// -ffp-contract=off, IEEE division/sqrt, denormals preserved), so host and
// device results are bit-identical to each other and to the reference CPU path.
//
//   Vector3 ops     Vector3.impl.h:9-33, 84-200
//   AABB            AABB.impl.h:12-16, 93-95, 116-130, 141-222
//   Triangle        Triangle.impl.h:185-208, 283-348
//   Shapes          Shapes.impl.cuh:200-312, 371-382
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#define DR_HD __host__ __device__ inline

namespace dr {

constexpr float kVecEps = 0.000005f;                 // VECTOR_EPSILON (Vector3.h:10)
constexpr float kInsideTol = 32 * kVecEps;           // Triangle.impl.h:299-301
constexpr float kOctEps = 256 * kVecEps;             // OCTREE_AABB_EPSILON (Octree.impl.cuh:3-4)

struct V3 { float x, y, z; };
struct Tri { V3 a, b, c; };
struct Box { V3 s, e; };

DR_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
DR_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
DR_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
DR_HD V3 operator*(V3 a, float f) { return v3(a.x * f, a.y * f, a.z * f); }
DR_HD V3 divf(V3 a, float f) { float fi = 1.0f / f; return v3(a.x * fi, a.y * fi, a.z * fi); }
DR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DR_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y); }
DR_HD float mag(V3 a) { return __builtin_sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
DR_HD bool eq(V3 a, V3 b) { return (a.x == b.x) && (a.y == b.y) && (a.z == b.z); }
// std::min / std::max semantics (the reference's unqualified min/max on the host)
DR_HD float smin(float a, float b) { return (b < a) ? b : a; }
DR_HD float smax(float a, float b) { return (a < b) ? b : a; }
// SHAPES_MIN / SHAPES_MAX (Shapes.impl.cuh:200-201)
DR_HD float shmin(float a, float b) { return (a < b) ? a : b; }
DR_HD float shmax(float a, float b) { return (a > b) ? a : b; }

DR_HD Tri tri(V3 a, V3 b, V3 c) { Tri t; t.a = a; t.b = b; t.c = c; return t; }

// AABB(s, e) with fix() (AABB.impl.h:12-16, 141-146)
DR_HD Box box(V3 s, V3 e) {
  float t;
  if (s.x > e.x) { t = s.x; s.x = e.x; e.x = t; }
  if (s.y > e.y) { t = s.y; s.y = e.y; e.y = t; }
  if (s.z > e.z) { t = s.z; s.z = e.z; e.z = t; }
  Box b; b.s = s; b.e = e; return b;
}
DR_HD bool contains(const Box &b, V3 v) {
  return v.x >= b.s.x && v.y >= b.s.y && v.z >= b.s.z && v.x <= b.e.x && v.y <= b.e.y && v.z <= b.e.z;
}
// AABB::castPreInversed != FLT_MAX  (AABB.impl.h:116-130 via Shapes.impl.cuh:144-147, clip=false)
DR_HD bool slab_hit(float sx, float sy, float sz, float ex, float ey, float ez, V3 o, V3 inv) {
  float ds = (sx - o.x) * inv.x;
  float de = (ex - o.x) * inv.x;
  float mn = smin(ds, de), mx = smax(ds, de);
  ds = (sy - o.y) * inv.y;
  de = (ey - o.y) * inv.y;
  mn = smax(mn, smin(ds, de));
  mx = smin(mx, smax(ds, de));
  ds = (sz - o.z) * inv.z;
  de = (ez - o.z) * inv.z;
  mn = smax(mn, smin(ds, de));
  mx = smin(mx, smax(ds, de));
  const float r = (mn > mx + kVecEps) ? FLT_MAX : mn;
  return r != FLT_MAX;
}

// slab_hit for a ray whose 1/d components are all finite.  Then no (s - o) * inv product is a
// NaN, min and max are exact (no rounding) and so associative, and the native min/max can
// differ from std::min/max only in the sign of a zero result, which neither comparison below
// sees: the decision is slab_hit's, bit for bit, with the IEEE min/max instructions.
DR_HD bool slab_hit_finite(float sx, float sy, float sz, float ex, float ey, float ez, V3 o, V3 inv) {
  const float x0 = (sx - o.x) * inv.x, x1 = (ex - o.x) * inv.x;
  const float y0 = (sy - o.y) * inv.y, y1 = (ey - o.y) * inv.y;
  const float z0 = (sz - o.z) * inv.z, z1 = (ez - o.z) * inv.z;
  const float mn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(x0, x1), __builtin_fminf(y0, y1)),
                                   __builtin_fminf(z0, z1));
  const float mx = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(x0, x1), __builtin_fmaxf(y0, y1)),
                                   __builtin_fmaxf(z0, z1));
  return !(mn > mx + kVecEps) && mn != FLT_MAX;
}

// Per-triangle constants of Triangle::containsVertex that do not depend on the ray
// (Triangle.impl.h:299-301 right-hand sides) and the plane normal (:335).
struct TriPre { V3 n; float thr0, thr1, thr2; bool degenerate; };
DR_HD TriPre tri_pre(const Tri &t) {
  TriPre p;
  p.n = cross(t.b - t.a, t.c - t.a);
  p.degenerate = eq(t.a, t.b) || eq(t.b, t.c) || eq(t.c, t.a);
  V3 ab = t.b - t.a, bc = t.c - t.b, ca = t.a - t.c;
  p.thr0 = -dot(ab, ca) / mag(ca);
  p.thr1 = -dot(bc, ab) / mag(ab);
  p.thr2 = -dot(ca, bc) / mag(bc);
  return p;
}

// ---- host-build helpers (Triangle sorting, triangle/box clipping) ----

// Triangle::sortByMases (Triangle.impl.h:185-199)
DR_HD void sort_by(Tri &t, float am, float bm, float cm) {
  V3 tmp;
  if (am > bm) {
    if (cm > am) { tmp = t.a; t.a = t.b; t.b = tmp; }
    else if (bm > cm) { tmp = t.a; t.a = t.c; t.c = tmp; }
    else { tmp = t.a; t.a = t.b; t.b = t.c; t.c = tmp; }
  } else {
    if (am > cm) { tmp = t.a; t.a = t.c; t.c = t.b; t.b = tmp; }
    else if (bm > cm) { tmp = t.b; t.b = t.c; t.c = tmp; }
  }
}
DR_HD float axis_of(V3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
DR_HD float axis_of_s(const Box &b, int axis) { return axis_of(b.s, axis); }
DR_HD float axis_of_e(const Box &b, int axis) { return axis_of(b.e, axis); }
// CROSS_POINT (AABB.impl.h:168, Shapes.impl.cuh:202)
DR_HD V3 cross_point(V3 from, V3 to, float fromV, float toV, float barrier) {
  return from + (to - from) * ((barrier - fromV) / (toV - fromV));
}
DR_HD float tri_surface(const Tri &t) { return mag(cross(t.b - t.a, t.c - t.a)) * 0.5f; }
DR_HD V3 tri_center(const Tri &t) { return divf((t.a + t.b) + t.c, 3.0f); }

}  // namespace dr
