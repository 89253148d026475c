// Internal helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <mutex>
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
  // references: the caller's (dr_scene_destroy drops it), one per dr_shading handle and per
  // batch request; the scene's memory is freed with the last one
  std::atomic<int> refs{1};
  int device = 0;
  dr::DevBuf groups, recs, urecs, lref, shade;
  dr::DevBuf faces;                        // per triangle: vertices a, b, c, then their normals (18 floats)
  std::vector<float> faces_host;
  dr::DevTree tree;
  int max_depth = 0;                       // depth the octree reached: traversal stack levels
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
  // Adjoint vertex records, one buffer per stream that has run dr_backward on this scene
  // (grown when a call needs more; calls on one stream are ordered, so they share it).
  struct StreamScratch {
    hipStream_t stream;
    dr::DevBuf buf;
  };
  std::mutex scratch_mu;
  std::vector<StreamScratch> scratch;
  int64_t bytes() const { return (int64_t)(groups.bytes + recs.bytes + urecs.bytes + lref.bytes + shade.bytes + faces.bytes + work.bytes); }
};

// One renderer's shading tables of a scene (dr_shading_create).
struct dr_shading_s {
  std::atomic<int> refs{1};
  dr_scene scene = nullptr;               // holds a reference
  int max_mat = -1;
  dr::DevBuf shade, faces;                // same layouts as the scene's own tables
};

namespace dr {
void scene_retain(dr_scene s);
void scene_release(dr_scene s);
void shading_retain(dr_shading h);
void shading_release(dr_shading h);
// The scene's vertex-record buffer for `stream`, at least `bytes` long.
int scene_scratch(dr_scene s, hipStream_t stream, size_t bytes, void **out);
// Device address of dr_backward_noalloc's scratch array on `device` (resolved once per device).
int noalloc_scratch_init(int device, float **out);
}  // namespace dr
