// Internal helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/dumbray.h"
#include "dr_octree_build.h"
#include "dr_traverse.h"

namespace dr {

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
  dr::DevBuf groups, recs, shade;
  dr::DevTree tree;
  int max_depth = 0;
  int ntris = 0;
  bool has_materials = false;
  int max_mat = -1;                        // largest material id in use
  // per triangle: n = (b-a) x (c-a) and the material id (bits of .w)
  std::vector<float4> shade_host;
  int64_t bytes() const { return (int64_t)(groups.bytes + recs.bytes + shade.bytes); }
};
