// dumbray.hpp — C++ host mirror of the reference's render-loop classes over the C ABI
// (include/dumbray.h).  Header-only; link with -ldumbray_hip and the HIP runtime.
//
// The shapes follow the reference so its render loop can switch over:
//   Vector3 / Triangle / BakedTriFace / Renderable<BakedTriFace>   (Vector3.h, Triangle.h,
//       BakedTriMesh.h:7-14, Material.cuh:140-161)
//   Octree<Renderable<BakedTriFace>>: reset / push / build / getNodeCount / getData / cast
//       (Octree.cuh:103-146); cast works on device batches, per-ray casts run inside kernels
//   BlockBasedFrameBuffer<8,16> (BlockBasedFrameBuffer.cuh:34-91), in device memory
//   DefaultPerspectiveLense / SimpleStochasticLense focal length and Camera transform
//       (DefaultPerspectiveLense.impl.cuh:5-9, Camera.cuh:12-24)
//   DumbRenderer::renderBlocksGPU (DumbRenderer.cuh:61-62) and Renderer::backward
// Errors throw dumbray::Error carrying dr_last_error().
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dumbray.h"

namespace dumbray {

struct Error : std::runtime_error {
  explicit Error(const std::string &what) : std::runtime_error(what + ": " + dr_last_error()) {}
};
inline void check(int rc, const char *what) {
  if (rc != DR_OK) throw Error(what);
}
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct Vector3 {
  float x, y, z;
};
struct Triangle {
  Vector3 a, b, c;
};
// BakedTriFace: vertex, normal and texture-coordinate triangles (only vert is used on the hot path).
struct BakedTriFace {
  Triangle vert, norm, tex;
};
template <typename Shape>
struct Renderable {
  Shape object;
  int materialId;
};

// Octree<Renderable<BakedTriFace>>: triangles are numbered in push order (hit id).
template <typename ElemType = Renderable<BakedTriFace>>
class Octree {
 public:
  explicit Octree(int splitThreshold = 8, int maxDepth = 24) : thresh_(splitThreshold), depth_(maxDepth) {}
  ~Octree() { release(); }
  Octree(const Octree &) = delete;
  Octree &operator=(const Octree &) = delete;

  void reset() {
    release();
    data_.clear();
  }
  void push(const ElemType &e) {
    release();
    data_.push_back(e);
  }
  void push(const std::vector<ElemType> &v) {
    release();
    data_.insert(data_.end(), v.begin(), v.end());
  }
  void build(int nthreads = 0) {
    release();
    std::vector<float> tris(9 * data_.size());
    for (size_t i = 0; i < data_.size(); i++) {
      const Triangle &t = data_[i].object.vert;
      const float v[9] = {t.a.x, t.a.y, t.a.z, t.b.x, t.b.y, t.b.z, t.c.x, t.c.y, t.c.z};
      for (int k = 0; k < 9; k++) tris[9 * i + k] = v[k];
    }
    check(dr_octree_build(tris.data(), (int)data_.size(), thresh_, depth_, nthreads, &tree_), "dr_octree_build");
    check(dr_scene_create(tree_, &scene_), "dr_scene_create");
    std::vector<int32_t> mat(data_.size());
    for (size_t i = 0; i < data_.size(); i++) mat[i] = data_[i].materialId;
    check(dr_scene_set_materials(scene_, mat.data(), (int)mat.size()), "dr_scene_set_materials");
  }
  int getNodeCount() const {
    int n = 0;
    check(dr_octree_info(tree_, &n, nullptr, nullptr, nullptr), "dr_octree_info");
    return n;
  }
  const std::vector<ElemType> &getData() const { return data_; }
  // Batch Octree::cast on the device: rays = n x (origin, direction).
  void cast(const float *rays_dev, int n, bool clipBackfaces, int32_t *ids_dev, float *t_dev,
            const int32_t *exclude_dev = nullptr, float *points_dev = nullptr, hipStream_t s = nullptr) const {
    check(dr_cast(scene_, rays_dev, n, clipBackfaces ? DR_CAST_CLIP_BACKFACES : 0u, exclude_dev, ids_dev, t_dev,
                  points_dev, (void *)s),
          "dr_cast");
  }
  dr_scene deviceScene() const { return scene_; }

 private:
  void release() {
    if (scene_) dr_scene_destroy(scene_);
    if (tree_) dr_octree_destroy(tree_);
    scene_ = nullptr;
    tree_ = nullptr;
  }
  int thresh_, depth_;
  std::vector<ElemType> data_;
  dr_octree tree_ = nullptr;
  dr_scene scene_ = nullptr;
};

// Device BlockBasedFrameBuffer<8,16>: RGBA fp32, block-major.
template <size_t blockW = 8, size_t blockH = 16>
class BlockBasedFrameBuffer {
  static_assert(blockW == 8 && blockH == 16, "the HIP renderer writes 8x16 blocks");

 public:
  BlockBasedFrameBuffer(int width, int height) { setResolution(width, height); }
  ~BlockBasedFrameBuffer() {
    if (data_) (void)hipFree(data_);
  }
  BlockBasedFrameBuffer(const BlockBasedFrameBuffer &) = delete;
  BlockBasedFrameBuffer &operator=(const BlockBasedFrameBuffer &) = delete;
  bool setResolution(int width, int height) {
    if (data_) (void)hipFree(data_);
    data_ = nullptr;
    w_ = width;
    h_ = height;
    hip_check(hipMalloc((void **)&data_, sizeof(float) * 4 * (size_t)getBlockSize() * getBlockCount()), "hipMalloc");
    hip_check(hipMemset(data_, 0, sizeof(float) * 4 * (size_t)getBlockSize() * getBlockCount()), "hipMemset");
    return true;
  }
  void getSize(int *width, int *height) const {
    *width = w_;
    *height = h_;
  }
  int getBlockSize() const { return (int)(blockW * blockH); }
  int getBlockCount() const { return ((w_ + (int)blockW - 1) / (int)blockW) * ((h_ + (int)blockH - 1) / (int)blockH); }
  float *deviceData() { return data_; }
  // BlockBasedFrameBuffer::pixelBlockLocation (BlockBasedFrameBuffer.impl.cuh:108-123)
  bool pixelBlockLocation(int x, int y, int *blockId, int *pixelId) const {
    if (x >= w_ || y >= h_) return false;
    const int nbw = (w_ + (int)blockW - 1) / (int)blockW;
    *blockId = (y / (int)blockH) * nbw + x / (int)blockW;
    *pixelId = (y % (int)blockH) * (int)blockW + x % (int)blockW;
    return true;
  }
  // Row-major RGBA copy of the frame (host).
  std::vector<float> toHostImage() const {
    std::vector<float> blocks(4 * (size_t)getBlockSize() * getBlockCount()), img(4 * (size_t)w_ * h_);
    hip_check(hipMemcpy(blocks.data(), data_, blocks.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int y = 0; y < h_; y++)
      for (int x = 0; x < w_; x++) {
        int b, p;
        pixelBlockLocation(x, y, &b, &p);
        for (int c = 0; c < 4; c++) img[4 * ((size_t)y * w_ + x) + c] = blocks[4 * ((size_t)b * getBlockSize() + p) + c];
      }
    return img;
  }

 private:
  int w_ = 0, h_ = 0;
  float *data_ = nullptr;
};
typedef BlockBasedFrameBuffer<8, 16> BlockBuffer;

// Camera = lens focal length + Transform (position and world axes).
struct Camera {
  Vector3 position, right, up, forward;
  float focal;   // DefaultPerspectiveLense: 0.5 / tan(fov / (2 * RADIAN))
  static float focalFromFov(float angleDegrees) {
    const float radian = (angleDegrees / (2 * 57.29577951308233f));
    return 0.5f / std::tan(radian);
  }
};

class DumbRenderer {
 public:
  enum BoxingMode { BOXING_MODE_HEIGHT_BASED = 0 };
  DumbRenderer(const Octree<> &geometry, const std::vector<dr_material> &materials,
               const std::vector<dr_light> &lights, const Camera &camera, int maxBounces = 2, int samplesPerPixelX = 1,
               int samplesPerPixelY = 1, bool ignoreBackfaces = true)
      : scene_(geometry.deviceScene()), camera_(camera), nmat_((int)materials.size()), nlights_((int)lights.size()) {
    hip_check(hipMalloc((void **)&mats_, sizeof(dr_material) * (materials.empty() ? 1 : materials.size())), "hipMalloc");
    hip_check(hipMalloc((void **)&lights_, sizeof(dr_light) * (lights.empty() ? 1 : lights.size())), "hipMalloc");
    if (!materials.empty())
      hip_check(hipMemcpy(mats_, materials.data(), sizeof(dr_material) * materials.size(), hipMemcpyHostToDevice), "copy");
    if (!lights.empty())
      hip_check(hipMemcpy(lights_, lights.data(), sizeof(dr_light) * lights.size(), hipMemcpyHostToDevice), "copy");
    setMaxBounces(maxBounces);
    setSamplesPerPixel(samplesPerPixelX, samplesPerPixelY);
    ignoreBackfaces_ = ignoreBackfaces;
  }
  ~DumbRenderer() {
    (void)hipFree(mats_);
    (void)hipFree(lights_);
  }
  DumbRenderer(const DumbRenderer &) = delete;
  DumbRenderer &operator=(const DumbRenderer &) = delete;

  void setMaxBounces(int b) { bounces_ = b < 64 ? b : 64; }        // DUMB_RENDERER_BOUNCE_LIMIT
  void setSamplesPerPixel(int x, int y) {
    fx_ = x <= 1 ? 1 : x;
    fy_ = y <= 1 ? 1 : y;
  }
  void setRussianRoulette(int startLayer, float survive) {
    rrStart_ = startLayer;
    rrSurvive_ = survive;
  }
  void setSeed(uint64_t seed) { seed_ = seed; }
  dr_material *deviceMaterials() { return mats_; }
  dr_light *deviceLights() { return lights_; }

  dr_render_params params(int width, int height, int iteration, uint32_t request = 0) const {
    dr_render_params p{};
    p.width = width;
    p.height = height;
    p.fsaa_x = fx_;
    p.fsaa_y = fy_;
    p.max_bounces = bounces_;
    p.rr_start = rrStart_;
    p.rr_survive = rrSurvive_;
    p.cull_backfaces = ignoreBackfaces_ ? 1 : 0;
    p.seed_lo = (uint32_t)seed_;
    p.seed_hi = (uint32_t)(seed_ >> 32);
    p.request = request;
    p.iteration = iteration;
    const float cam[13] = {camera_.position.x, camera_.position.y, camera_.position.z, camera_.right.x,
                           camera_.right.y,    camera_.right.z,    camera_.up.x,       camera_.up.y,
                           camera_.up.z,       camera_.forward.x,  camera_.forward.y,  camera_.forward.z,
                           camera_.focal};
    for (int k = 0; k < 13; k++) p.cam[k] = cam[k];
    p.materials = mats_;
    p.lights = lights_;
    p.nmat = nmat_;
    p.nlights = nlights_;
    return p;
  }
  // DumbRenderer::renderBlocksGPU: blocks [startBlock, endBlock) of `iteration` into the device frame.
  bool renderBlocksGPU(BlockBuffer &frame, int startBlock, int endBlock, int iteration, hipStream_t stream = nullptr) {
    int w, h;
    frame.getSize(&w, &h);
    const dr_render_params p = params(w, h, iteration);
    return dr_render_blocks(scene_, &p, frame.deviceData(), startBlock, endBlock, (void *)stream) == DR_OK;
  }
  // Renderer::backward: adds dL/dalbedo (double[nmat*3]) and dL/demission (double[nlights*3]), device buffers.
  void backward(BlockBuffer &dlDframe, int iteration, double *gAlbedoDev, double *gEmissionDev,
                hipStream_t stream = nullptr) {
    int w, h;
    dlDframe.getSize(&w, &h);
    const dr_render_params p = params(w, h, iteration);
    check(dr_backward(scene_, &p, dlDframe.deviceData(), gAlbedoDev, gEmissionDev, (void *)stream), "dr_backward");
  }

 private:
  dr_scene scene_;
  Camera camera_;
  dr_material *mats_ = nullptr;
  dr_light *lights_ = nullptr;
  int nmat_, nlights_;
  int bounces_ = 2, fx_ = 1, fy_ = 1, rrStart_ = 1 << 30;
  float rrSurvive_ = 1.0f;
  bool ignoreBackfaces_ = true;
  uint64_t seed_ = 0x5EED;
};

}  // namespace dumbray
