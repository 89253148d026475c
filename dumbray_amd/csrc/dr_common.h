// Internal helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <string>
#include <vector>
#include "../../include/dumbray.h"
#include "dr_octree_build.h"
#include "dr_traverse.h"

namespace dr {

constexpr int kWorkSlots = 1024;         // cast work-counter pairs per scene

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
#define DR_HIP_CHECK(expr)                                                                            \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return ::dr::fail(DR_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Device buffer owned by a scene.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};
int dev_upload(DevBuf &b, const void *host, size_t bytes);
void dev_free(DevBuf &b);

}  // namespace dr

struct dr_octree_s {
  dr::OctreeHost host;
};

struct dr_scene_s {
  int device = 0;
  dr::DevBuf groups, recs, urecs, lref, shade;
  dr::DevBuf faces;                        // per triangle: vertices a, b, c, then their normals (18 floats)
  std::vector<float> faces_host;
  dr::DevTree tree;
  int max_depth = 0;
  int ntris = 0;
  bool has_materials = false;
  int max_mat = -1;                        // largest material id in use
  // per triangle: n = (b-a) x (c-a) and the material id (bits of .w)
  std::vector<float4> shade_host;
  // Work counters of the cast kernels: kWorkSlots pairs (next ray, waves done), zero
  // between launches; each launch takes the next pair, so launches in flight at the
  // same time (up to kWorkSlots of them) use distinct counters.
  dr::DevBuf work;
  std::atomic<uint32_t> work_slot{0};
  int num_cu = 1;
  int64_t bytes() const { return (int64_t)(groups.bytes + recs.bytes + urecs.bytes + lref.bytes + shade.bytes + faces.bytes + work.bytes); }
};
