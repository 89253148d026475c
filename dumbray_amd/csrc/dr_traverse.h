// Device-side octree closest-hit traversal (Octree::cast, Octree.impl.cuh:174-212,
// castInLeaf :584-612, Triangle::cast Triangle.impl.h:334-348) over the
// structure-of-arrays group layout built by flatten_octree().
//
// Order of work is the reference's: children of an expanded node are visited
// in the order canonical ^ i (i = 0..7), a leaf that accepts a hit ends the
// cast, and inside a leaf the first triangle with the strictly smallest t wins.
// The box tests of all 8 children are evaluated together when a node is
// expanded (they are pure, so this only changes when, not whether, they run).
// Children whose subtree leaves all lie outside the ray's forward octant are
// dropped: no hit point o + d*t with t >= 0 can be inside their boxes, so the
// reference would visit them and reject every hit.
#pragma once
#include "dr_geom.h"
#include "dr_octree_build.h"

namespace dr {

struct DevTree {
  const float4 *__restrict__ groups;   // kGroupF4 float4 per group
  const float4 *__restrict__ recs;     // 4 float4 per triangle record, one per leaf entry
  const float4 *__restrict__ urecs;    // the same records once per distinct triangle
  const int *__restrict__ lref;        // leaf entry -> record in urecs
  float rsx, rsy, rsz, rex, rey, rez;  // root box
  int root_is_leaf, root_start, root_count;
  float margin;                        // prune margin of the octant test (see the group layout)
};

struct Hit {
  int id;        // -1 miss
  float t;
  V3 p;
};

// Work counters of one cast (measurement launches only; nullptr elsewhere).
struct CastCounters {
  int expand;    // groups expanded (196-B reads: 8 boxes + non-empty mask)
  int leaves;    // leaves whose triangles were tested (+ leaf box and link/count reads)
  int plane;     // triangle records whose plane test ran (first 32 B)
  int inside;    // records that reached the inside test (second 32 B)
  int live;      // children of expanded groups with a non-empty leaf below (octant tests run)
  int slab;      // children whose slab test ran (live and not behind the ray's origin)
};

// A ray prepared for traversal.
struct TRay {
  V3 o, d, inv;
  int canon;     // canonicalOrder (Octree.impl.cuh:181-183)
  V3 hi, lo;     // prune thresholds: a child is behind if max.k < hi.k or min.k > lo.k (see set_prune)
};
__device__ __forceinline__ TRay make_tray(V3 o, V3 d) {
  TRay r;
  r.o = o;
  r.d = d;
  r.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int c = (d.z < 0) ? 1 : 0;
  if (d.y < 0) c += 2;
  if (d.x < 0) c += 4;
  r.canon = c;
  return r;
}

// v rounded down (up) by at least one ulp of v: below (above) every real number that
// rounds to v.
__device__ __forceinline__ float below(float v) { return v - (__builtin_fabsf(v) * 2.4e-7f + 1e-37f); }
__device__ __forceinline__ float above(float v) { return v + (__builtin_fabsf(v) * 2.4e-7f + 1e-37f); }
// Prune thresholds of the forward-octant test for margin M.  On an axis with d.k >= 0
// a child whose box max satisfies max.k < hi.k = below(o.k - M) has max.k + M < o.k in
// exact arithmetic, so every non-empty leaf below it (within the box widened by M) lies
// behind o along k; symmetrically lo.k = above(o.k + M) for d.k <= 0.  Axes that do not
// apply get -inf / +inf.
__device__ __forceinline__ void set_prune(TRay &r, float M) {
  const float inf = __builtin_inff();
  r.hi = v3(r.d.x >= 0 ? below(r.o.x - M) : -inf, r.d.y >= 0 ? below(r.o.y - M) : -inf,
            r.d.z >= 0 ? below(r.o.z - M) : -inf);
  r.lo = v3(r.d.x <= 0 ? above(r.o.x + M) : inf, r.d.y <= 0 ? above(r.o.y + M) : inf,
            r.d.z <= 0 ? above(r.o.z + M) : inf);
}

// One edge of Triangle::containsVertex (Triangle.impl.h:299-301): fl(fl(D / fl(sqrt(S))) + tol) >= thr
// with D = dot(edge, p - vertex), S = dot(p - vertex, p - vertex), correctly rounded division
// and square root.
__device__ __forceinline__ bool edge_exact(float D, float S, float thr) {
  return D / __builtin_sqrtf(S) + kInsideTol >= thr;
}
// The same decision from the hardware reciprocal square root where it is certain: 1 = pass,
// 0 = fail, 2 = too close to call.  q = D * rsq(S) is within about 2 ulp of the correctly
// rounded quotient (rsq within 1 ulp, one product rounding, the quotient's two roundings), and
// the sum with tol adds at most one ulp more.  An edge is decided only when q + tol lies more
// than kEdgeMargin * (|q| + tol), about 34 ulp, from thr; S outside [1e-30, 1e30] (where
// denormals or overflow could enter) is left to the exact form.
constexpr float kEdgeMargin = 4e-6f;
__device__ __forceinline__ int edge_fast(float D, float S, float thr) {
  if (!(S >= 1e-30f && S <= 1e30f)) return 2;
  const float q = D * __builtin_amdgcn_rsqf(S);
  const float r = q + kInsideTol;
  const float m = kEdgeMargin * (__builtin_fabsf(q) + kInsideTol);
  return (r - m >= thr) ? 1 : ((r + m < thr) ? 0 : 2);
}

// Triangle::containsVertex minus the degenerate test (degenerate triangles have no record).
// The three edge conditions are ANDed; each is decided by edge_fast where it is certain and by
// the exact form otherwise, so the result is the reference's bit for bit.
__device__ __forceinline__ bool inside(V3 p, V3 a, V3 b, V3 c, float thr0, float thr1, float thr2) {
  if (eq(p, a) || eq(p, b) || eq(p, c)) return true;
  const V3 ab = b - a, bc = c - b, ca = a - c;
  const V3 ax = p - a, bx = p - b, cx = p - c;
  const float D0 = dot(ab, ax), S0 = dot(ax, ax), D1 = dot(bc, bx), S1 = dot(bx, bx), D2 = dot(ca, cx),
              S2 = dot(cx, cx);
  const int e0 = edge_fast(D0, S0, thr0), e1 = edge_fast(D1, S1, thr1), e2 = edge_fast(D2, S2, thr2);
  if (e0 == 0 || e1 == 0 || e2 == 0) return false;
  if ((e0 & e1 & e2) == 1) return true;
  return (e0 == 1 || edge_exact(D0, S0, thr0)) && (e1 == 1 || edge_exact(D1, S1, thr1)) &&
         (e2 == 1 || edge_exact(D2, S2, thr2));
}

// Box of child k of group g (the leaf-acceptance box).
__device__ __forceinline__ Box child_box(const float4 *__restrict__ G, int g, int k) {
  const float *f = (const float *)(G + kGroupF4 * (size_t)g);
  Box b;
  b.s = v3(f[k], f[8 + k], f[16 + k]);
  b.e = v3(f[24 + k], f[32 + k], f[40 + k]);
  return b;
}

// Traversal stack.  A level holds two words, [2*level][lane] and [2*level+1][lane]:
//   lo = group << 8 | mask of children still to visit (bit i = child canon ^ i)
//   hi = first interior child's group << 8 | interior mask (bit k = child k)
// so popping an interior child needs no memory read.
constexpr int kStackWordsPerLevel = 2;

// 8 child tests of group g -> stack entry.  A child is kept if it has a non-empty
// leaf below it, its box widened by the prune margin meets the forward octant of
// the ray (the thresholds of set_prune), and the reference slab test on its box passes.
template <bool kFinite>
__device__ __forceinline__ void group_entry(const float4 *__restrict__ G, int g, const TRay &r, uint32_t &lo,
                                            uint32_t &hi, CastCounters *cc = nullptr) {
  const float4 *q = G + kGroupF4 * (size_t)g;
  // children with no non-empty leaf below them are dropped before their box tests run
  const uint2 head = *(const uint2 *)(q + kGroupMask / 4);
  uint32_t m = 0;
  // all 192 B of child boxes requested at once, before any child is tested: the loads of the
  // second four children no longer wait behind the tests of the first four (incoherent cast
  // 2.41 -> 2.34 ms, 1080p render 26.3 -> 26.0 ms)
  float4 qv[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) qv[k] = q[k];
#pragma unroll
  for (int h = 0; h < 2; ++h) {       // children in two halves of 4
    const float4 sx = qv[0 + h], sy = qv[2 + h], sz = qv[4 + h], ex = qv[6 + h], ey = qv[8 + h], ez = qv[10 + h];
    const float a_sx[4] = {sx.x, sx.y, sx.z, sx.w}, a_sy[4] = {sy.x, sy.y, sy.z, sy.w},
                a_sz[4] = {sz.x, sz.y, sz.z, sz.w}, a_ex[4] = {ex.x, ex.y, ex.z, ex.w},
                a_ey[4] = {ey.x, ey.y, ey.z, ey.w}, a_ez[4] = {ez.x, ez.y, ez.z, ez.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // forward octant of o: coordinate k of o + d*t (t >= 0) is >= o.k if d.k >= 0, <= o.k if d.k <= 0
      const bool live = (head.x >> (4 * h + j)) & 1u;
      const bool behind = !live || (a_ex[j] < r.hi.x) || (a_sx[j] > r.lo.x) || (a_ey[j] < r.hi.y) ||
                          (a_sy[j] > r.lo.y) || (a_ez[j] < r.hi.z) || (a_sz[j] > r.lo.z);
      if (cc) { cc->live += live ? 1 : 0; cc->slab += behind ? 0 : 1; }
      const bool hit =
          !behind && (kFinite ? slab_hit_finite(a_sx[j], a_sy[j], a_sz[j], a_ex[j], a_ey[j], a_ez[j], r.o, r.inv)
                              : slab_hit(a_sx[j], a_sy[j], a_sz[j], a_ex[j], a_ey[j], a_ez[j], r.o, r.inv));
      m |= hit ? (1u << (4 * h + j)) : 0u;
    }
  }
  const int c = r.canon;
  if (c & 1) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
  if (c & 2) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
  if (c & 4) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
  lo = ((uint32_t)g << 8) | m;
  hi = (head.y << 8) | ((head.x >> 8) & 0xffu);
}

// Initial traversal state.  Returns 0 = miss, 1 = traversal pending, 2 = root is a leaf.
// The root's group (group 0, the root's children) is expanded by the first walk_step,
// with the expansions of the other lanes, rather than here: a new ray is set up in the
// divergent code that hands out work, where an expansion would run with few lanes.
__device__ __forceinline__ int tree_begin(const DevTree &T, const TRay &r, int &sp) {
  if (T.root_is_leaf) return 2;
  if (!slab_hit(T.rsx, T.rsy, T.rsz, T.rex, T.rey, T.rez, r.o, r.inv)) return 0;
  sp = -1;
  return 1;
}

// One lane's closest-hit query, advanced in bounded steps so that a kernel can
// interleave many of them (and batch their leaf tests) across a wave.
enum : int { kWalkTrace = 1, kWalkLeaf = 2, kWalkDone = 3 };
struct Walk {
  TRay ray;
  int ex;              // validateNotSameAsObject's object, or -1
  int sp, state;
  int lg, lk, llink, lcnt;   // pending leaf: child lk of group lg (lg < 0: the root)
  bool hit;            // result once state == kWalkDone
  Hit h;
};

__device__ __forceinline__ void walk_begin(const DevTree &T, Walk &w, V3 o, V3 d, int exclude, uint32_t *stk,
                                           int stride, CastCounters *cc) {
  w.ray = make_tray(o, d);
  set_prune(w.ray, T.margin);
  w.ex = exclude;
  w.hit = false;
  const int st = tree_begin(T, w.ray, w.sp);
  if (st == 0) {
    w.state = kWalkDone;
  } else if (st == 2) {
    w.state = kWalkLeaf;
    w.lg = -1; w.llink = T.root_start; w.lcnt = T.root_count;
  } else {
    w.state = kWalkTrace;
  }
}

// Pops one child: expands an interior child, or holds a leaf child as pending;
// an exhausted stack ends the query with a miss.  A walk's first step (sp == -1 after
// walk_begin) expands the root's group.
// kFastSlab: expand with the NaN-free slab test whenever no expanding lane of the wave has an
// infinite 1/d (kernels with registers to spare for the second copy of the child test).
template <bool kFastSlab = true>
__device__ __forceinline__ void walk_step(const DevTree &T, Walk &w, uint32_t *stk, int stride, CastCounters *cc) {
  int child = 0;                      // group to expand; 0 (the root's) on a walk's first step
  if (w.sp >= 0) {
    uint32_t f = 0;
    while (w.sp >= 0) {
      f = stk[2 * w.sp * stride];
      if (f & 0xffu) break;
      --w.sp;
    }
    if (w.sp < 0) { w.state = kWalkDone; return; }
    const int i = __builtin_ctz(f & 0xffu);
    stk[2 * w.sp * stride] = f & ~(1u << i);
    const int k = w.ray.canon ^ i;
    const uint32_t hi = stk[(2 * w.sp + 1) * stride];
    const uint32_t im = hi & 0xffu;
    if (!((im >> k) & 1u)) {
      const int *Gi = (const int *)T.groups;
      const size_t base = kGroupWords * (size_t)(f >> 8);
      w.lg = (int)(f >> 8);
      w.lk = k;
      w.llink = Gi[base + kGroupLink + k];
      w.lcnt = Gi[base + kGroupCount + k];
      w.state = kWalkLeaf;
      return;
    }
    child = (int)(hi >> 8) + __builtin_popcount(im & ((1u << k) - 1u));
  }
  uint32_t lo2, hi2;
  // the NaN-free form of the slab test when no expanding lane of the wave has an infinite 1/d
  const bool finite = __builtin_isfinite(w.ray.inv.x) && __builtin_isfinite(w.ray.inv.y) &&
                      __builtin_isfinite(w.ray.inv.z);
  if (kFastSlab && __ballot(!finite) == 0) group_entry<true>(T.groups, child, w.ray, lo2, hi2, cc);
  else group_entry<false>(T.groups, child, w.ray, lo2, hi2, cc);
  ++w.sp;
  stk[2 * w.sp * stride] = lo2;
  stk[(2 * w.sp + 1) * stride] = hi2;
  if (cc) cc->expand++;
}

// Inclusive prefix sum of v over the 64 lanes of the wave (every lane active): DPP row
// shifts by 1, 2, 4 and 8 (a lane without a source adds 0) scan each row of 16 lanes, and
// each row adds the totals of the rows before it, read from lanes 15, 31 and 47 into scalar
// registers.  `total` gets the sum over all lanes.  Unlike a scan over __shfl_up, it needs
// no per-lane address registers, which the compiler would otherwise keep live across the
// caller's loop.
template <int kCtrl>
__device__ __forceinline__ int dpp_shr_or_zero(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xf, 0xf, false);
}
__device__ __forceinline__ int wave_incl_scan(int v, int &total) {
  v += dpp_shr_or_zero<0x111>(v);   // row_shr:1
  v += dpp_shr_or_zero<0x112>(v);   // row_shr:2
  v += dpp_shr_or_zero<0x114>(v);   // row_shr:4
  v += dpp_shr_or_zero<0x118>(v);   // row_shr:8
  const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31),
            r2 = __builtin_amdgcn_readlane(v, 47), r3 = __builtin_amdgcn_readlane(v, 63);
  total = (r0 + r1) + (r2 + r3);
  const int row = (int)__lane_id() >> 4;
  return v + (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
}

// castInLeaf for every pending lane of the wave at once, with all 64 lanes sharing
// the records.  Must be called by every lane of the wave (wave-uniform control flow);
// `pending` marks the lanes whose leaf is tested; `best` (and `ins`, counting the
// records that reach the inside test, for measurement launches) hold one slot per
// lane of the wave in LDS.
//   * The records of the pending leaves are numbered 0..total-1 in lane order (wave
//     prefix sum of the counts); lane j tests records j, j + 64, ...
//   * A record's owner is the last lane whose first record number is <= the record's
//     (counted over the prefix sums in two levels, below); its ray, exclusion and first
//     record come from the owner with __shfl.
//   * A record that passes Triangle::cast with t < FLT_MAX offers the 64-bit key
//     (t bits << 32 | record's position in its leaf) to the owner's slot in `best`
//     (LDS, one uint64 per lane) with an atomic min.  t > 0, so the keys order by t
//     first and by position second: the owner gets the first record with the smallest
//     t, which is what the sequential strict '<' scan keeps (Octree.impl.cuh:595-606).
//   * The owner then applies the leaf-box acceptance.
// kIndexed: read the records through lref from urecs (a third of the footprint on
// scenes whose triangles sit in several leaves; one more dependent load per record, issued
// one round ahead).
// The next round's owners and record indices are found during the current round (incoherent
// cast 2.31 -> 2.12 ms with the lref read ahead; 1080p render 25.6 -> 24.1 ms and the
// 64-request batch 247 -> 228 ms with the owner search ahead).
template <bool kIndexed = false>
__device__ __forceinline__ void walk_leaf_wave(const DevTree &T, Walk &w, bool pending, bool clip,
                                               unsigned long long *best, int *ins, CastCounters *cc) {
  const int lane = (int)__lane_id();
  const int cnt = pending ? w.lcnt : 0;
  int total;
  const int incl = wave_incl_scan(cnt, total);
  const int start = incl - cnt;
  if (pending) __hip_atomic_store(&best[lane], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  if (cc && pending) __hip_atomic_store(&ins[lane], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  const int link = pending ? w.llink : 0;
  // The owner of record j is the last lane whose first record number is <= j (the numbers
  // never decrease from lane to lane), i.e. the count of lanes 1..63 whose number is <= j.
  // Level 1 counts among lanes 8, 16, ..., 56, whose numbers sit in scalar registers for the
  // whole batch; level 2 among the 7 lanes after the multiple of 8 found, fetched together.
  int coarse[7];
#pragma unroll
  for (int m = 0; m < 7; m++) coarse[m] = __builtin_amdgcn_readlane(start, 8 * (m + 1));
  // record j's owner (wave-uniform call)
  auto owner_of = [&](int j) {
    int b = 0;
#pragma unroll
    for (int m = 0; m < 7; m++) b += coarse[m] <= j ? 8 : 0;
    int owner = b;
#pragma unroll
    for (int i = 1; i < 8; i++) owner += __shfl(start, b + i) <= j ? 1 : 0;
    return owner;
  };
  // tests record j (of owner `owner`; its 64 B in r0..r3) for its owner; every lane takes part
  auto test_record = [&](int j, int owner, float4 r0, float4 r1, float4 r2, float4 r3) {
    const V3 o = v3(__shfl(w.ray.o.x, owner), __shfl(w.ray.o.y, owner), __shfl(w.ray.o.z, owner));
    const V3 d = v3(__shfl(w.ray.d.x, owner), __shfl(w.ray.d.y, owner), __shfl(w.ray.d.z, owner));
    const int pos = j - __shfl(start, owner);
    const int ex = __shfl(w.ex, owner);
    if (j < total) {
      const V3 a = v3(r0.x, r0.y, r0.z), n = v3(r0.w, r1.x, r1.y);
      const float dp = dot(a - o, n);
      const float dd = dot(d, n);
      const float t = dp / dd;
      if (!(clip && (dp < -kVecEps)) && !(dp * dd <= 0) && (t < FLT_MAX)) {
        if (cc) __hip_atomic_fetch_add(&ins[owner], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (__float_as_int(r3.w) != ex) {
          const V3 p = o + d * t;
          const V3 b = v3(r1.z, r1.w, r2.x), c = v3(r2.y, r2.z, r2.w);
          if (inside(p, a, b, c, r3.x, r3.y, r3.z)) {
            const unsigned long long key =
                ((unsigned long long)__float_as_uint(t) << 32) | (unsigned long long)(unsigned)pos;
            __hip_atomic_fetch_min(&best[owner], key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
          }
        }
      }
    }
  };
  // The index of a lane's record in the next round (its owner, and with kIndexed the lref entry)
  // is found while the current round's records are fetched and tested, so a round's record
  // loads start at once.
  const float4 *R = kIndexed ? T.urecs : T.recs;
  int own = owner_of(lane);
  int kn = 0;
  {
    const int kel0 = __shfl(link, own) + lane - __shfl(start, own);   // every lane takes part
    if (lane < total) kn = kIndexed ? T.lref[kel0] : kel0;
  }
  for (int base = 0; base < total; base += 64) {
    const int j = base + lane;
    const int owner = own;
    const int k = j < total ? kn : 0;      // (lanes past the last record read record 0)
    const float4 r0 = R[4 * k + 0], r1 = R[4 * k + 1], r2 = R[4 * k + 2], r3 = R[4 * k + 3];
    if (base + 64 < total) {
      own = owner_of(j + 64);
      const int kel = __shfl(link, own) + (j + 64) - __shfl(start, own);
      if (j + 64 < total) kn = kIndexed ? T.lref[kel] : kel;
    }
    test_record(j, owner, r0, r1, r2, r3);
  }
  if (!pending) return;
  if (cc) {
    cc->leaves++;
    cc->plane += cnt;
    cc->inside += __hip_atomic_load(&ins[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
  const unsigned long long key = __hip_atomic_load(&best[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  bool hit = false;
  if (key != ~0ull) {
    const float t = __uint_as_float((unsigned)(key >> 32));
    const int k = w.llink + (int)(key & 0xffffffffu);
    const V3 p = w.ray.o + w.ray.d * t;
    // the record's id is read beside the leaf box, before the box test that decides whether
    // it is needed
    const int rid = kIndexed ? __float_as_int(T.urecs[4 * T.lref[k] + 3].w) : __float_as_int(T.recs[4 * k + 3].w);
    Box lb;
    if (w.lg < 0) { lb.s = v3(T.rsx, T.rsy, T.rsz); lb.e = v3(T.rex, T.rey, T.rez); }
    else lb = child_box(T.groups, w.lg, w.lk);
    if (contains(lb, p)) {
      hit = true;
      w.h.id = rid;
      w.h.t = t;
      w.h.p = p;
    }
  }
  w.hit = hit;
  w.state = (hit || w.lg < 0) ? kWalkDone : kWalkTrace;
}

}  // namespace dr
