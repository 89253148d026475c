// Host octree builder: Octree<Renderable<BakedTriFace>>::push(all) + build()
// (Octree.impl.cuh:97-129, 339-533) re-designed as a parallel builder (large nodes
// split by the whole thread team, the subtrees below them built one per thread)
// whose output is the device layout the traversal kernel reads.
//
// The tree it produces (node boxes, leaf lists in stored order, child order)
// is the reference's tree; only node *indices* differ, exactly as in the
// reference's own thread-pool build (scheduling permutes indices, the data
// fixes everything else).  Splits are independent once a node's triangle list
// is final, so every subtree is built by whichever thread takes it.
#include "dr_octree_build.h"

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>
#include <cmath>
#include <cstring>

namespace dr {
namespace {

// ---------------------------------------------------------------- clipping
// The clip works on a small fixed buffer: three slab passes turn one triangle
// into at most 3 * 3 * 3 = 27 pieces (Shapes.impl.cuh:255-266).
struct ClipBuf { Tri v[32]; int n; };

// One SHAPES_PUSH_INTERSECTIONS pass (Shapes.impl.cuh:203-253).
void clip_pass(ClipBuf &front, ClipBuf &back, int axis, float s, float e) {
  for (int i = 0; i < front.n; i++) {
    Tri &t = front.v[i];
    sort_by(t, axis_of(t.a, axis), axis_of(t.b, axis), axis_of(t.c, axis));
    const float av = axis_of(t.a, axis), bv = axis_of(t.b, axis), cv = axis_of(t.c, axis);
    if (cv < s) continue;
    else if (av > e) continue;
    else if (av <= s) {
      V3 asc = cross_point(t.a, t.c, av, cv, s);
      if (bv <= s) {
        V3 bsc = cross_point(t.b, t.c, bv, cv, s);
        if (cv <= e) back.v[back.n++] = tri(asc, bsc, t.c);
        else {
          V3 bec = cross_point(t.b, t.c, bv, cv, e);
          back.v[back.n++] = tri(bsc, bec, asc);
          V3 aec = cross_point(t.a, t.c, av, cv, e);
          back.v[back.n++] = tri(asc, bec, aec);
        }
      } else if (bv <= e) {
        if (cv <= e) {
          back.v[back.n++] = tri(asc, t.b, t.c);
          V3 asb = cross_point(t.a, t.b, av, bv, s);
          back.v[back.n++] = tri(asc, asb, t.b);
        } else {
          V3 asb = cross_point(t.a, t.b, av, bv, s);
          V3 bec = cross_point(t.b, t.c, bv, cv, e);
          back.v[back.n++] = tri(asb, t.b, bec);
          back.v[back.n++] = tri(asc, asb, bec);
          V3 aec = cross_point(t.a, t.c, av, cv, e);
          back.v[back.n++] = tri(asc, bec, aec);
        }
      } else {
        V3 asb = cross_point(t.a, t.b, av, bv, s);
        V3 aeb = cross_point(t.a, t.b, av, bv, e);
        back.v[back.n++] = tri(asc, asb, aeb);
        V3 aec = cross_point(t.a, t.c, av, cv, e);
        back.v[back.n++] = tri(asc, aeb, aec);
      }
    } else {
      if (cv <= e) back.v[back.n++] = t;
      else {
        V3 aec = cross_point(t.a, t.c, av, cv, e);
        if (bv <= e) {
          V3 bec = cross_point(t.b, t.c, bv, cv, e);
          back.v[back.n++] = tri(t.a, t.b, bec);
          back.v[back.n++] = tri(t.a, aec, bec);
        } else {
          V3 aeb = cross_point(t.a, t.b, av, bv, e);
          back.v[back.n++] = tri(t.a, aeb, aec);
        }
      }
    }
  }
  front.n = 0;
}
// x, then y, then z (SHAPES_GET_INTERSECTING_TRIANGLES); pieces end in `out`.
void clip_to_box(const Box &b, const Tri &t, ClipBuf &out, ClipBuf &tmp) {
  out.n = 0;
  tmp.n = 1;
  tmp.v[0] = t;
  clip_pass(tmp, out, 0, b.s.x, b.e.x);
  clip_pass(out, tmp, 1, b.s.y, b.e.y);
  clip_pass(tmp, out, 2, b.s.z, b.e.z);
}
// intersectionCenter<AABB, Triangle> (Shapes.impl.cuh:293-312)
V3 isect_center(const Box &b, const Tri &t, ClipBuf &l, ClipBuf &tmp) {
  clip_to_box(b, t, l, tmp);
  if (l.n > 0) {
    float size = 0.0f;
    V3 sum = v3(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < l.n; i++) {
      const float surface = (tri_surface(l.v[i]) + 0.0001f);
      sum = sum + tri_center(l.v[i]) * surface;
      size += surface;
    }
    return divf(sum, size);
  }
  return divf(b.s + b.e, 2.0f);
}
// intersectionBounds<AABB, Triangle> (Shapes.impl.cuh:272-291)
Box isect_bounds(const Box &b, const Tri &t, ClipBuf &l, ClipBuf &tmp) {
  clip_to_box(b, t, l, tmp);
  if (l.n <= 0) return b;
  V3 mn = b.e, mx = b.s;
  for (int i = 0; i < l.n; i++) {
    const Tri &q = l.v[i];
    V3 nmx = v3(shmax(shmax(shmax(mx.x, q.a.x), q.b.x), q.c.x), shmax(shmax(shmax(mx.y, q.a.y), q.b.y), q.c.y),
                shmax(shmax(shmax(mx.z, q.a.z), q.b.z), q.c.z));
    V3 nmn = v3(shmin(shmin(shmin(mn.x, q.a.x), q.b.x), q.c.x), shmin(shmin(shmin(mn.y, q.a.y), q.b.y), q.c.y),
                shmin(shmin(shmin(mn.z, q.a.z), q.b.z), q.c.z));
    mx = nmx;
    mn = nmn;
  }
  return box(mn, mx);
}
// AABB::intersectsTri<dimm> (AABB.impl.h:150-222); recursion axis order z, x, y.
bool box_hits_tri(const Box &b, Tri t, int dimm) {
  if (dimm >= 3) return true;
  const int axis = (dimm == 0) ? 2 : (dimm == 1 ? 0 : 1);
  sort_by(t, axis_of(t.a, axis), axis_of(t.b, axis), axis_of(t.c, axis));
  const float av = axis_of(t.a, axis), bv = axis_of(t.b, axis), cv = axis_of(t.c, axis);
  const float s = axis_of_s(b, axis), e = axis_of_e(b, axis);
  const int nd = dimm + 1;
  if (cv < s) return false;
  if (av > e) return false;
  if (av <= s) {
    V3 asc = cross_point(t.a, t.c, av, cv, s);
    if (bv <= s) {
      V3 bsc = cross_point(t.b, t.c, bv, cv, s);
      if (cv <= e) return box_hits_tri(b, tri(asc, bsc, t.c), nd);
      V3 bec = cross_point(t.b, t.c, bv, cv, e);
      if (box_hits_tri(b, tri(bsc, bec, asc), nd)) return true;
      V3 aec = cross_point(t.a, t.c, av, cv, e);
      return box_hits_tri(b, tri(asc, bec, aec), nd);
    } else if (bv <= e) {
      if (cv <= e) {
        if (box_hits_tri(b, tri(asc, t.b, t.c), nd)) return true;
        V3 asb = cross_point(t.a, t.b, av, bv, s);
        return box_hits_tri(b, tri(asc, asb, t.b), nd);
      }
      V3 asb = cross_point(t.a, t.b, av, bv, s);
      V3 bec = cross_point(t.b, t.c, bv, cv, e);
      if (box_hits_tri(b, tri(asb, t.b, bec), nd)) return true;
      if (box_hits_tri(b, tri(asc, asb, bec), nd)) return true;
      V3 aec = cross_point(t.a, t.c, av, cv, e);
      return box_hits_tri(b, tri(asc, bec, aec), nd);
    } else {
      V3 asb = cross_point(t.a, t.b, av, bv, s);
      V3 aeb = cross_point(t.a, t.b, av, bv, e);
      if (box_hits_tri(b, tri(asc, asb, aeb), nd)) return true;
      V3 aec = cross_point(t.a, t.c, av, cv, e);
      return box_hits_tri(b, tri(asc, aeb, aec), nd);
    }
  }
  if (cv <= e) return box_hits_tri(b, t, nd);
  V3 aec = cross_point(t.a, t.c, av, cv, e);
  if (bv <= e) {
    V3 bec = cross_point(t.b, t.c, bv, cv, e);
    if (box_hits_tri(b, tri(t.a, t.b, bec), nd)) return true;
    return box_hits_tri(b, tri(t.a, aec, bec), nd);
  }
  V3 aeb = cross_point(t.a, t.b, av, bv, e);
  return box_hits_tri(b, tri(t.a, aeb, aec), nd);
}
// boundingBox<Triangle> (Shapes.impl.cuh:371-382)
Box tri_bbox(const Tri &t) {
  return box(v3(shmin(shmin(t.a.x, t.b.x), t.c.x), shmin(shmin(t.a.y, t.b.y), t.c.y), shmin(shmin(t.a.z, t.b.z), t.c.z)),
             v3(shmax(shmax(t.a.x, t.b.x), t.c.x), shmax(shmax(t.a.y, t.b.y), t.c.y), shmax(shmax(t.a.z, t.b.z), t.c.z)));
}

// ---------------------------------------------------------------- node pool
// Nodes live in fixed chunks that never move, so workers can hold references
// while others append.  Children are appended 8 at a time (one group).
struct BNode {
  Box box;
  int32_t child = -1;                 // first of 8 contiguous children, -1 = leaf
  std::vector<int32_t> list;          // triangle ids in stored order (leaves)
};
class NodePool {
 public:
  static constexpr int kShift = 13, kChunk = 1 << kShift, kMaxChunks = 1 << 17;
  // A run of nodes one thread allocates from without locking, refilled a chunk at a time.
  struct Cursor {
    int32_t next = 0, end = 0;
  };
  NodePool() : chunks_(new std::unique_ptr<BNode[]>[kMaxChunks]) {}
  BNode &operator[](int32_t i) { return chunks_[i >> kShift][i & (kChunk - 1)]; }
  int32_t alloc(int k, Cursor &cur) {   // k consecutive nodes (1 or 8), never straddling a chunk
    if (cur.next + k > cur.end) {
      std::lock_guard<std::mutex> g(mu_);
      const int c = used_;
      if (c >= kMaxChunks) return -1;
      chunks_[c].reset(new BNode[kChunk]);
      used_ = c + 1;
      cur.next = c << kShift;
      cur.end = cur.next + kChunk;
    }
    const int32_t first = cur.next;
    cur.next += k;
    return first;
  }
  // upper bound of the node indices handed out (unused slots are never reached from the root)
  int32_t size() const { return used_ << kShift; }
 private:
  std::unique_ptr<std::unique_ptr<BNode[]>[]> chunks_;
  std::mutex mu_;
  int used_ = 0;
};

// Runs fn(begin, end) over [0, n) in contiguous chunks of at least `grain` elements on
// up to `threads` threads (the caller's thread takes the first chunk).  Used inside the split of a large node, where the
// per-triangle clipping dominates; every result is written per element and combined
// afterwards in element order, so the tree does not depend on the chunking.
template <class Fn>
void parallel_chunks(int n, int threads, Fn fn, int grain = 4096) {
  threads = std::max(1, std::min(threads, n / grain + 1));
  if (threads == 1) { fn(0, n); return; }
  std::vector<std::thread> team;
  const int step = (n + threads - 1) / threads;
  for (int t = 1; t < threads; t++) {
    const int b = t * step, e = std::min(n, b + step);
    if (b < e) team.emplace_back([=] { fn(b, e); });
  }
  fn(0, std::min(n, step));
  for (auto &th : team) th.join();
}

struct Builder {
  const Tri *tris;
  int thresh, maxdepth;
  V3 Ev;
  int threads = 1;                    // team size for the split of a large node
  static constexpr int kBigNode = 16384;
  NodePool pool;
  std::atomic<bool> failed{false};

  // splitAABB (Octree.impl.cuh:423-434)
  void split_aabb(const Box &b, V3 c, Box *r) const {
    const V3 s = b.s, e = b.e;
    const float E = kOctEps;
    r[0] = box(s, c + Ev);
    r[1] = box(v3(s.x, s.y, c.z - E), v3(c.x + E, c.y + E, e.z));
    r[2] = box(v3(s.x, c.y - E, s.z), v3(c.x + E, e.y, c.z + E));
    r[3] = box(v3(s.x, c.y - E, c.z - E), v3(c.x + E, e.y, e.z));
    r[4] = box(v3(c.x - E, s.y, s.z), v3(e.x, c.y + E, c.z + E));
    r[5] = box(v3(c.x - E, s.y, c.z - E), v3(e.x, c.y + E, e.z));
    r[6] = box(v3(c.x - E, c.y - E, s.z), v3(e.x, e.y, c.z + E));
    r[7] = box(c - Ev, e);
  }
  // reduceNode(index, true) on a leaf (Octree.impl.cuh:489-533, leaf branch); min / max
  // are exact, so a large leaf's bounds may be gathered in chunks and merged.
  void reduce_leaf(BNode &nd, ClipBuf &l, ClipBuf &tmp) const {
    const int n = (int)nd.list.size();
    auto gather = [&](int b, int e, ClipBuf &cl, ClipBuf &ct, V3 &start, V3 &end) {
      for (int j = b; j < e; j++) {
        const Box ob = isect_bounds(nd.box, tris[nd.list[j]], cl, ct);
        start = v3(smin(start.x, ob.s.x), smin(start.y, ob.s.y), smin(start.z, ob.s.z));
        end = v3(smax(end.x, ob.e.x), smax(end.y, ob.e.y), smax(end.z, ob.e.z));
      }
    };
    V3 start = nd.box.e, end = nd.box.s;
    if (n < kBigNode || threads <= 1) {
      gather(0, n, l, tmp, start, end);
    } else {
      std::mutex mu;
      parallel_chunks(n, threads, [&](int b, int e) {
        ClipBuf cl, ct;
        V3 s0 = nd.box.e, e0 = nd.box.s;
        gather(b, e, cl, ct, s0, e0);
        std::lock_guard<std::mutex> g(mu);
        start = v3(smin(start.x, s0.x), smin(start.y, s0.y), smin(start.z, s0.z));
        end = v3(smax(end.x, e0.x), smax(end.y, e0.y), smax(end.z, e0.z));
      });
    }
    finish_reduce(nd, start, end);
  }
  void finish_reduce(BNode &nd, V3 start, V3 end) const {
    start = start - Ev;
    end = end + Ev;
    const V3 bs = nd.box.s, be = nd.box.e;
    start = v3(smax(start.x, bs.x), smax(start.y, bs.y), smax(start.z, bs.z));
    end = v3(smin(end.x, be.x), smin(end.y, be.y), smin(end.z, be.z));
    nd.box = box(start, end);
  }
  // split (Octree.impl.cuh:339-421) of one node; appends its children to `out`.
  void split(int32_t idx, int depth, ClipBuf &l, ClipBuf &tmp, std::vector<uint8_t> &cache,
             std::vector<std::pair<int32_t, int>> &out, NodePool::Cursor &cur) {
    BNode &nd = pool[idx];
    const int n = (int)nd.list.size();
    if (depth >= maxdepth || n < thresh) { reduce_leaf(nd, l, tmp); return; }
    V3 center = v3(0, 0, 0);
    const Box bounds = nd.box;
    Box sub[8];
    int insiders[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    cache.assign(n, 0);
    if (n < kBigNode || threads <= 1) {
      for (int i = 0; i < n; i++) center = center + isect_center(bounds, tris[nd.list[i]], l, tmp);
      center = divf(center, (float)n);
      split_aabb(bounds, center, sub);
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < n; j++)
          if (box_hits_tri(sub[i], tris[nd.list[j]], 0)) { insiders[i]++; cache[j] |= (uint8_t)(1u << i); }
    } else {
      // the clipped centres per element in parallel, summed in element order as the loop above
      std::vector<V3> part(n);
      parallel_chunks(n, threads, [&](int b, int e) {
        ClipBuf cl, ct;
        for (int j = b; j < e; j++) part[j] = isect_center(bounds, tris[nd.list[j]], cl, ct);
      });
      for (int i = 0; i < n; i++) center = center + part[i];
      center = divf(center, (float)n);
      split_aabb(bounds, center, sub);
      std::mutex mu;
      parallel_chunks(n, threads, [&](int b, int e) {
        int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = b; j < e; j++)
          for (int i = 0; i < 8; i++)
            if (box_hits_tri(sub[i], tris[nd.list[j]], 0)) { cnt[i]++; cache[j] |= (uint8_t)(1u << i); }
        std::lock_guard<std::mutex> g(mu);
        for (int i = 0; i < 8; i++) insiders[i] += cnt[i];
      });
    }
    // splittingMakesSence with the intersection cache (Octree.impl.cuh:437-475)
    int full = 0;
    float load = 0.0f;
    for (int i = 0; i < 8; i++) {
      if (insiders[i] >= n) full++;
      load += ((float)insiders[i]) / ((float)n);
    }
    if (full >= 1 || !(load < smin(((float)n) / ((float)thresh), 6.0f))) { reduce_leaf(nd, l, tmp); return; }
    // splitNode + distribution (Octree.impl.cuh:388-404, 477-486)
    const int32_t c0 = pool.alloc(8, cur);
    if (c0 < 0) { failed = true; return; }
    for (int i = 0; i < 8; i++) pool[c0 + i].box = sub[i];
    if (n < kBigNode || threads <= 1) {
      for (int j = 0; j < n; j++) {
        const int32_t id = nd.list[j];
        for (int i = 0; i < 8; i++) if (cache[j] & (1u << i)) pool[c0 + i].list.push_back(id);
      }
    } else {
      // one child per task, each list still in the parent's order
      parallel_chunks(8, threads, [&](int b, int e) {
        for (int i = b; i < e; i++) {
          std::vector<int32_t> &cl = pool[c0 + i].list;
          cl.reserve(insiders[i]);
          for (int j = 0; j < n; j++) if (cache[j] & (1u << i)) cl.push_back(nd.list[j]);
        }
      }, 1);
    }
    std::vector<int32_t>().swap(nd.list);
    nd.child = c0;
    for (int i = 7; i >= 0; i--) out.emplace_back(c0 + i, depth + 1);
  }
  // The subtree below one node, depth first on the calling thread (the order of the
  // reference's recursion; node indices do not affect the tree).
  void build_subtree(std::pair<int32_t, int> task, NodePool::Cursor &cur) {
    ClipBuf l, tmp;
    std::vector<uint8_t> cache;
    std::vector<std::pair<int32_t, int>> stack{task}, out;
    while (!stack.empty() && !failed) {
      const auto t = stack.back();
      stack.pop_back();
      out.clear();
      split(t.first, t.second, l, tmp, cache, out, cur);
      stack.insert(stack.end(), out.begin(), out.end());
    }
  }
  // Large nodes first, one at a time with the whole team inside each split; then the
  // subtrees below them, handed out to the threads largest first.
  void run(int32_t root) {
    NodePool::Cursor cur;
    std::vector<std::pair<int32_t, int>> frontier{{root, 0}}, small, out;
    ClipBuf l, tmp;
    std::vector<uint8_t> cache;
    while (!frontier.empty() && !failed) {
      std::vector<std::pair<int32_t, int>> next;
      for (const auto &t : frontier) {
        if ((int)pool[t.first].list.size() < kBigNode) { small.push_back(t); continue; }
        out.clear();
        split(t.first, t.second, l, tmp, cache, out, cur);
        next.insert(next.end(), out.begin(), out.end());
      }
      frontier.swap(next);
    }
    std::stable_sort(small.begin(), small.end(), [&](const std::pair<int32_t, int> &a, const std::pair<int32_t, int> &b) {
      return pool[a.first].list.size() > pool[b.first].list.size();
    });
    std::atomic<size_t> at{0};
    auto work = [&] {
      NodePool::Cursor c;
      for (size_t i = at++; i < small.size() && !failed; i = at++) build_subtree(small[i], c);
    };
    std::vector<std::thread> team;
    for (int i = 1; i < threads; i++) team.emplace_back(work);
    work();
    for (auto &th : team) th.join();
  }
  // reduceNode(0, false) (Octree.impl.cuh:489-533): interior boxes from their children
  void reduce_interior(int32_t idx) {
    BNode &nd = pool[idx];
    if (nd.child < 0) return;
    for (int i = 0; i < 8; i++) reduce_interior(nd.child + i);
    V3 start = nd.box.e, end = nd.box.s;
    for (int i = 0; i < 8; i++) {
      const Box b = pool[nd.child + i].box;
      start = v3(smin(start.x, b.s.x), smin(start.y, b.s.y), smin(start.z, b.s.z));
      end = v3(smax(end.x, b.e.x), smax(end.y, b.e.y), smax(end.z, b.e.z));
    }
    finish_reduce(nd, start, end);
  }
};

}  // namespace

int build_octree(const float *tris9, int ntris, int split_threshold, int max_depth, int nthreads, OctreeHost &out) {
  if (ntris < 0 || split_threshold < 1 || max_depth < 0 || max_depth > kMaxTreeDepth) return -1;
  out = OctreeHost();
  out.split_threshold = split_threshold;
  out.max_depth = max_depth;
  out.tris.resize(ntris);
  for (int i = 0; i < ntris; i++) {
    const float *p = tris9 + 9 * (size_t)i;
    out.tris[i] = tri(v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]), v3(p[6], p[7], p[8]));
  }
  Builder B;
  B.tris = out.tris.data();
  B.thresh = split_threshold;
  B.maxdepth = max_depth;
  B.Ev = v3(kVecEps, kVecEps, kVecEps) * 256;        // EPSILON_VECTOR * 256
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  nthreads = std::min(nthreads, 64);
  B.threads = nthreads;
  NodePool::Cursor rc0;
  const int32_t root = B.pool.alloc(1, rc0);
  BNode &r = B.pool[root];
  r.box = box(v3(0, 0, 0), v3(0, 0, 0));
  // push (Octree.impl.cuh:97-111): first element sets the root box, later ones union box +- eps
  r.list.resize(ntris);
  for (int i = 0; i < ntris; i++) {
    const Box bb = tri_bbox(out.tris[i]);
    if (i == 0) r.box = bb;
    else {
      const V3 os = r.box.s, oe = r.box.e;
      const V3 bs = bb.s - B.Ev, be = bb.e + B.Ev;
      r.box = box(v3(smin(os.x, bs.x), smin(os.y, bs.y), smin(os.z, bs.z)),
                  v3(smax(oe.x, be.x), smax(oe.y, be.y), smax(oe.z, be.z)));
    }
    r.list[i] = i;
  }
  B.run(root);
  if (B.failed) return -2;
  B.reduce_interior(root);

  // Export in breadth-first order so the node numbering is deterministic and the
  // interior children of a node own consecutive child blocks.
  const int32_t total = B.pool.size();
  out.bounds.clear();
  out.children.clear();
  out.list_off.clear();
  out.list.clear();
  out.bounds.reserve(total);
  std::vector<int32_t> remap(total, -1);
  // queue over groups: node -> new index; children stay contiguous
  std::vector<int32_t> order;
  order.reserve(total);
  order.push_back(root);
  remap[root] = 0;
  for (size_t k = 0; k < order.size(); k++) {
    const BNode &nd = B.pool[order[k]];
    if (nd.child >= 0)
      for (int i = 0; i < 8; i++) { remap[nd.child + i] = (int32_t)order.size(); order.push_back(nd.child + i); }
  }
  out.list_off.push_back(0);
  for (int32_t old : order) {
    const BNode &nd = B.pool[old];
    out.bounds.push_back(nd.box);
    out.children.push_back(nd.child >= 0 ? remap[nd.child] : -1);
    out.list.insert(out.list.end(), nd.list.begin(), nd.list.end());
    out.list_off.push_back((int32_t)out.list.size());
  }
  return 0;
}

}  // namespace dr

namespace dr {

static void put_rec(std::vector<float> &recs, const Tri &t, int32_t id) {
  const TriPre p = tri_pre(t);
  float w[kRecWords] = {t.a.x, t.a.y, t.a.z, p.n.x, p.n.y, p.n.z, t.b.x, t.b.y,
                        t.b.z, t.c.x, t.c.y, t.c.z, p.thr0, p.thr1, p.thr2, 0.0f};
  memcpy(&w[15], &id, 4);
  recs.insert(recs.end(), w, w + kRecWords);
}

int flatten_octree(const OctreeHost &t, OctreeFlat &out) {
  out = OctreeFlat();
  const int nn = (int)t.bounds.size();
  if (nn < 1) return -1;
  out.root = t.bounds[0];
  // deepest node the build reached (breadth-first order: parents precede their children);
  // a walk holds one stack level per interior node on its path, so this many levels suffice
  {
    std::vector<int> depth(nn, 0);
    for (int i = 0; i < nn; i++)
      if (t.children[i] >= 0)
        for (int c = t.children[i]; c < t.children[i] + 8; c++) {
          depth[c] = depth[i] + 1;
          out.max_depth = std::max(out.max_depth, depth[c]);
        }
  }
  auto nondegenerate = [&](int32_t id) {
    const Tri &tr = t.tris[id];
    return !(eq(tr.a, tr.b) || eq(tr.b, tr.c) || eq(tr.c, tr.a));    // degenerate never hits (Triangle.impl.h:292)
  };
  std::vector<int32_t> urec_of(t.tris.size(), -1);
  auto leaf_records = [&](int node, int32_t &start, int32_t &count) {
    start = (int32_t)(out.recs.size() / kRecWords);
    count = 0;
    for (int k = t.list_off[node]; k < t.list_off[node + 1]; k++) {
      const int32_t id = t.list[k];
      if (!nondegenerate(id)) continue;
      put_rec(out.recs, t.tris[id], id);
      if (urec_of[id] < 0) {
        urec_of[id] = (int32_t)(out.urecs.size() / kRecWords);
        put_rec(out.urecs, t.tris[id], id);
      }
      out.lref.push_back(urec_of[id]);
      count++;
    }
  };
  if (t.children[0] < 0) {
    out.root_is_leaf = 1;
    leaf_records(0, out.root_start, out.root_count);
    out.nrec = (int)(out.recs.size() / kRecWords);
    return 0;
  }
  out.ngroups = (nn - 1) / 8;
  if ((nn - 1) % 8 != 0 || out.ngroups >= (1 << 24)) return -1;
  // subtree leaf bounds, children before parents (breadth-first order reversed)
  const float inf = __builtin_inff();
  std::vector<Box> slb(nn);
  for (int i = nn - 1; i >= 0; i--) {
    Box b;
    b.s = v3(inf, inf, inf);
    b.e = v3(-inf, -inf, -inf);
    if (t.children[i] < 0) {
      bool any = false;
      for (int k = t.list_off[i]; k < t.list_off[i + 1] && !any; k++) any = nondegenerate(t.list[k]);
      if (any) b = t.bounds[i];
    } else {
      for (int c = t.children[i]; c < t.children[i] + 8; c++) {
        const Box &q = slb[c];
        b.s = v3(std::min(b.s.x, q.s.x), std::min(b.s.y, q.s.y), std::min(b.s.z, q.s.z));
        b.e = v3(std::max(b.e.x, q.e.x), std::max(b.e.y, q.e.y), std::max(b.e.z, q.e.z));
      }
    }
    slb[i] = b;
  }
  // prune margin: how far any non-empty leaf box reaches beyond the box of any of its
  // ancestors (evaluated in double), plus a guard of 4 ulps of the largest coordinate so
  // that the fp32 sum box + margin is never below the leaf's face.
  double poke = 0.0, maxabs = 0.0;
  for (int i = 1; i < nn; i++) {
    const Box &b = t.bounds[i], &l = slb[i];
    if (!(l.s.x <= l.e.x)) continue;                   // no non-empty leaf below
    const double d[6] = {(double)l.e.x - b.e.x, (double)l.e.y - b.e.y, (double)l.e.z - b.e.z,
                         (double)b.s.x - l.s.x, (double)b.s.y - l.s.y, (double)b.s.z - l.s.z};
    for (double v : d) poke = std::max(poke, v);
    const double c[6] = {b.s.x, b.s.y, b.s.z, b.e.x, b.e.y, b.e.z};
    for (double v : c) maxabs = std::max(maxabs, std::fabs(v));
  }
  out.prune_margin = (float)(poke + 4.0 * maxabs * 1.1920928955078125e-07 + 1e-30);
  out.groups.assign((size_t)out.ngroups * kGroupWords, 0.0f);
  for (int g = 0; g < out.ngroups; g++) {
    float *G = &out.groups[(size_t)g * kGroupWords];
    uint32_t nonempty = 0, interior = 0;
    int32_t first_child = 0;
    for (int k = 0; k < 8; k++) {
      const int node = 1 + 8 * g + k;
      const Box &b = t.bounds[node];
      G[0 + k] = b.s.x; G[8 + k] = b.s.y; G[16 + k] = b.s.z;
      G[24 + k] = b.e.x; G[32 + k] = b.e.y; G[40 + k] = b.e.z;
      if (slb[node].s.x <= slb[node].e.x) nonempty |= 1u << k;
      int32_t link, count;
      if (t.children[node] >= 0) {
        link = (t.children[node] - 1) / 8;
        count = -1;
        const int rank = __builtin_popcount(interior);
        if (rank == 0) first_child = link;
        else if (link != first_child + rank) return -1;   // not breadth-first
        interior |= 1u << k;
      } else {
        leaf_records(node, link, count);
      }
      memcpy(&G[kGroupLink + k], &link, 4);
      memcpy(&G[kGroupCount + k], &count, 4);
    }
    const uint32_t head = nonempty | (interior << 8);
    memcpy(&G[kGroupMask], &head, 4);
    memcpy(&G[kGroupBase], &first_child, 4);
  }
  out.nrec = (int)(out.recs.size() / kRecWords);
  return 0;
}

}  // namespace dr
