// Device-side octree closest-hit traversal (Octree::cast, Octree.impl.cuh:174-212,
// castInLeaf :584-612, Triangle::cast Triangle.impl.h:334-348) over the
// structure-of-arrays group layout built by flatten_octree().
//
// Order of work is the reference's: children of an expanded node are visited
// in the order canonical ^ i (i = 0..7), a leaf that accepts a hit ends the
// cast, and inside a leaf the first triangle with the strictly smallest t wins.
// The box tests of all 8 children are evaluated together when a node is
// expanded (they are pure, so this only changes when, not whether, they run).
#pragma once
#include "dr_geom.h"

namespace dr {

struct DevTree {
  const float4 *__restrict__ groups;   // 16 float4 per group (kGroupWords)
  const float4 *__restrict__ recs;     // 4 float4 per triangle record (kRecWords)
  float rsx, rsy, rsz, rex, rey, rez;  // root box
  int root_is_leaf, root_start, root_count;
  int pad;
};

struct Hit {
  int id;        // -1 miss
  float t;
  V3 p;
};

// Triangle::containsVertex minus the degenerate test (degenerate triangles have no record)
__device__ __forceinline__ bool inside(V3 p, V3 a, V3 b, V3 c, float thr0, float thr1, float thr2) {
  if (eq(p, a) || eq(p, b) || eq(p, c)) return true;
  const V3 ab = b - a, bc = c - b, ca = a - c;
  const V3 ax = p - a, bx = p - b, cx = p - c;
  return (dot(ab, ax) / mag(ax) + kInsideTol >= thr0) && (dot(bc, bx) / mag(bx) + kInsideTol >= thr1) &&
         (dot(ca, cx) / mag(cx) + kInsideTol >= thr2);
}

// castInLeaf: closest accepted triangle among `count` records; accepted iff its hit
// point lies inside the leaf box.  exclude = validateNotSameAsObject's object (or -1).
__device__ __forceinline__ bool leaf_cast(const float4 *__restrict__ R, int start, int count, const Box &lb, V3 o,
                                          V3 d, bool clip, int exclude, Hit &h) {
  float best = FLT_MAX;
  V3 bp = v3(0.f, 0.f, 0.f);
  int bid = -1;
  for (int k = start; k < start + count; ++k) {
    const float4 r0 = R[4 * k + 0], r1 = R[4 * k + 1], r2 = R[4 * k + 2];
    const V3 a = v3(r0.x, r0.y, r0.z), n = v3(r2.y, r2.z, r2.w);
    const float dp = dot(a - o, n);
    if (clip && (dp < -kVecEps)) continue;
    const float dd = dot(d, n);
    if (dp * dd <= 0) continue;
    const float t = dp / dd;
    if (!(t < best)) continue;                       // could not replace the current best
    const float4 r3 = R[4 * k + 3];
    const int id = __float_as_int(r3.w);
    if (id == exclude) continue;
    const V3 p = o + d * t;
    const V3 b = v3(r0.w, r1.x, r1.y), c = v3(r1.z, r1.w, r2.x);
    if (inside(p, a, b, c, r3.x, r3.y, r3.z)) { best = t; bp = p; bid = id; }
  }
  if (best != FLT_MAX && contains(lb, bp)) { h.id = bid; h.t = best; h.p = bp; return true; }
  return false;
}

// 8 child box tests of group g -> mask in visiting order (bit i = child canonical ^ i),
// empty leaves removed (castInLeaf on them returns false).
__device__ __forceinline__ uint32_t group_mask(const float4 *__restrict__ G, int g, V3 o, V3 inv, int canon) {
  const float4 *q = G + 16 * (size_t)g;
  float sx[8], sy[8], sz[8], ex[8], ey[8], ez[8];
  int cnt[8];
  *(float4 *)&sx[0] = q[0];  *(float4 *)&sx[4] = q[1];
  *(float4 *)&sy[0] = q[2];  *(float4 *)&sy[4] = q[3];
  *(float4 *)&sz[0] = q[4];  *(float4 *)&sz[4] = q[5];
  *(float4 *)&ex[0] = q[6];  *(float4 *)&ex[4] = q[7];
  *(float4 *)&ey[0] = q[8];  *(float4 *)&ey[4] = q[9];
  *(float4 *)&ez[0] = q[10]; *(float4 *)&ez[4] = q[11];
  *(float4 *)&cnt[0] = q[14]; *(float4 *)&cnt[4] = q[15];
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    m |= ((cnt[k] != 0) && slab_hit(sx[k], sy[k], sz[k], ex[k], ey[k], ez[k], o, inv)) ? (1u << k) : 0u;
  if (canon & 1) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
  if (canon & 2) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
  if (canon & 4) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
  return m;
}

// Octree::cast.  stk points at this lane's first stack word; consecutive levels are
// `stride` words apart (LDS, [level][lane] so a wave touches 64 consecutive words).
__device__ __forceinline__ bool tree_cast(const DevTree &T, V3 o, V3 d, bool clip, int exclude, uint32_t *stk,
                                          int stride, Hit &h) {
  if (T.root_is_leaf) {
    Box rb; rb.s = v3(T.rsx, T.rsy, T.rsz); rb.e = v3(T.rex, T.rey, T.rez);
    return leaf_cast(T.recs, T.root_start, T.root_count, rb, o, d, clip, exclude, h);
  }
  const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  if (!slab_hit(T.rsx, T.rsy, T.rsz, T.rex, T.rey, T.rez, o, inv)) return false;
  int canon = (d.z < 0) ? 1 : 0;
  if (d.y < 0) canon += 2;
  if (d.x < 0) canon += 4;
  const float4 *__restrict__ G = T.groups;
  const int *__restrict__ Gi = (const int *)G;
  const float *__restrict__ Gf = (const float *)G;
  int sp = 0;
  stk[0] = group_mask(G, 0, o, inv, canon);           // (group << 8) | mask, group 0 = root's children
  while (sp >= 0) {
    const uint32_t f = stk[sp * stride];
    const uint32_t m = f & 0xffu;
    if (m == 0) { --sp; continue; }
    const int i = __builtin_ctz(m);
    stk[sp * stride] = f & ~(1u << i);
    const int g = (int)(f >> 8);
    const int k = canon ^ i;
    const size_t base = 64 * (size_t)g;
    const int link = Gi[base + 48 + k];
    const int cnt = Gi[base + 56 + k];
    if (cnt < 0) {
      ++sp;
      stk[sp * stride] = ((uint32_t)link << 8) | group_mask(G, link, o, inv, canon);
    } else {
      Box lb;
      lb.s = v3(Gf[base + k], Gf[base + 8 + k], Gf[base + 16 + k]);
      lb.e = v3(Gf[base + 24 + k], Gf[base + 32 + k], Gf[base + 40 + k]);
      if (leaf_cast(T.recs, link, cnt, lb, o, d, clip, exclude, h)) return true;
    }
  }
  return false;
}

}  // namespace dr
