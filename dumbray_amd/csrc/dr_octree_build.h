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
// A "group" is the 8 contiguous children of one interior node, stored as 64
// 32-bit words (256 B, one aligned pair of 128-B lines):
//   [0..7]   min.x of child 0..7     [24..31] max.x
//   [8..15]  min.y                   [32..39] max.y
//   [16..23] min.z                   [40..47] max.z
//   [48..55] link: interior -> child group index; leaf -> first triangle record
//   [56..63] count: -1 interior, 0 empty leaf (or only degenerate triangles), >0 leaf records
// Child k uses the reference's numbering (bit0 = z-high, bit1 = y-high, bit2 = x-high).
constexpr int kGroupWords = 64;
// A triangle record is 16 words (64 B): a.xyz b.xyz c.xyz n.xyz thr0 thr1 thr2 id,
// n = (b-a) x (c-a) and thr* = the ray-independent right-hand sides of the
// containsVertex test.  Degenerate triangles (always a miss) get no record.
constexpr int kRecWords = 16;

struct OctreeFlat {
  std::vector<float> groups;               // ngroups * 64 words (ints stored bitwise)
  std::vector<float> recs;                 // nrec * 16 words
  Box root;
  int root_is_leaf = 0, root_start = 0, root_count = 0;
  int ngroups = 0, nrec = 0, max_depth = 0;
};
int flatten_octree(const OctreeHost &t, OctreeFlat &out);

}  // namespace dr
