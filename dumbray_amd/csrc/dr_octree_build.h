// Host-side octree (the reference's Octree<Renderable<BakedTriFace>> after
// push + build) and its flattening into the device traversal layout.
#pragma once
#include <vector>
#include "dr_geom.h"

namespace dr {

constexpr int kMaxTreeDepth = 24;          // OCTREE_MAX_DEPTH (Octree.cuh:13); traversal stack bound

struct OctreeHost {
  int split_threshold = 8, max_depth = 24;
  std::vector<Tri> tris;                   // push order = hit id
  // nodes in breadth-first order; the 8 children of a node are contiguous
  std::vector<Box> bounds;
  std::vector<int32_t> children;           // first child, -1 = leaf
  std::vector<int32_t> list_off, list;     // leaf triangle lists (stored order)
};

// Builds the octree over tris9 (ntris x 9 floats).  nthreads <= 0: all cores.
// Returns 0, -1 for bad arguments, -2 if the node pool overflowed.
int build_octree(const float *tris9, int ntris, int split_threshold, int max_depth, int nthreads, OctreeHost &out);

// ---------------------------------------------------------------- device layout
// A "group" is the 8 contiguous children of one interior node, stored as 72
// 32-bit words (288 B):
//   [0..47]  reference boxes: min.x[8] min.y[8] min.z[8] max.x[8] max.y[8] max.z[8]
//   [48]     bits 0-7: child k has a non-empty leaf at or below it;
//            bits 8-15: child k is interior
//   [49]     group of the first interior child.  Interior children get consecutive
//            groups (breadth-first order), so interior child k's group is this
//            plus the number of interior children before k.
//   [50..55] unused
//   [56..63] link: interior -> child group index; leaf -> first triangle record
//   [64..71] count: -1 interior, 0 empty leaf (or only degenerate triangles), >0 records
// Child k uses the reference's numbering (bit0 = z-high, bit1 = y-high, bit2 = x-high).
// Traversal drops a child whose box, widened by the tree's prune margin, lies
// outside the ray's forward octant: a leaf accepts a hit only if the hit point
// o + d*t (t >= 0) is inside its box, that point never leaves the closed octant
// of o that d points into, and every non-empty leaf box lies inside each of its
// ancestors' boxes widened by the margin (computed exactly when flattening).
constexpr int kGroupWords = 72;
constexpr int kGroupF4 = kGroupWords / 4;
constexpr int kGroupMask = 48, kGroupBase = 49, kGroupLink = 56, kGroupCount = 64;
// A triangle record is 16 words (64 B):
//   a.x a.y a.z n.x | n.y n.z b.x b.y | b.z c.x c.y c.z | thr0 thr1 thr2 id
// n = (b-a) x (c-a) and thr* = the ray-independent right-hand sides of the
// containsVertex test; the plane test reads only the first 32 B.
// Degenerate triangles (always a miss) get no record.  Records exist in two layouts:
// one per leaf entry in leaf order (read by the render kernels) and one per distinct
// triangle with a leaf-entry -> record index (read by the closest-hit kernel, whose
// random rays touch the whole scene: the compact layout fits the last-level cache).
constexpr int kRecWords = 16;

struct OctreeFlat {
  std::vector<float> groups;               // ngroups * kGroupWords words (ints stored bitwise)
  std::vector<float> recs;                 // nrec * 16 words, one per leaf entry (leaf order)
  std::vector<float> urecs;                // the same records once per distinct triangle
  std::vector<int32_t> lref;               // leaf entry -> its record in urecs
  Box root;
  int root_is_leaf = 0, root_start = 0, root_count = 0;
  int ngroups = 0, nrec = 0;
  int max_depth = 0;                       // depth of the deepest node (root: 0), not the build limit
  float prune_margin = 0.0f;               // see the group layout above
};
int flatten_octree(const OctreeHost &t, OctreeFlat &out);

}  // namespace dr
