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
//   DumbRenderContext::buildFromFile (scene files), BufferedRenderProcess, Images::saveBufferPNG
// Errors throw dumbray::Error carrying dr_last_error().
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <ctime>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
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
    // BakedTriFace::norm, when the faces carry vertex normals (all-zero normals mean flat shading)
    std::vector<float> nrm(9 * data_.size());
    bool any = false;
    for (size_t i = 0; i < data_.size(); i++) {
      const Triangle &n = data_[i].object.norm;
      const float v[9] = {n.a.x, n.a.y, n.a.z, n.b.x, n.b.y, n.b.z, n.c.x, n.c.y, n.c.z};
      for (int k = 0; k < 9; k++) { nrm[9 * i + k] = v[k]; any = any || v[k] != 0.0f; }
    }
    if (any) check(dr_scene_set_normals(scene_, nrm.data(), (int)data_.size()), "dr_scene_set_normals");
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

// Render plugin constructors: each does what the reference plugin's constructor does, in
// fp32, so the device table holds the values the reference object would.  Textures are
// not supported (every ColoredTexture is its colour).
namespace Plugins {
inline dr_material material(int32_t kind, Vector3 albedo) {
  dr_material m{};
  m.kind = kind;
  m.albedo[0] = albedo.x; m.albedo[1] = albedo.y; m.albedo[2] = albedo.z;
  return m;
}
inline dr_material lambert(Vector3 albedo) { return material(DR_LAMBERT, albedo); }
inline dr_material mirror(Vector3 reflectance) { return material(DR_MIRROR, reflectance); }
inline dr_material metal(Vector3 f0) { return material(DR_METAL, f0); }
// DumbBasedShader (DumbBasedShader.impl.cuh:5-16): metal clamped to [0, 1], diffuse *= 1 - metal.
inline dr_material dumbBasedShader(Vector3 fresnel = {0.2f, 0.2f, 0.2f}, float specular = 8.0f,
                                   Vector3 diffuse = {0.2f, 0.2f, 0.2f}, float metalness = 0.5f) {
  dr_material m = material(DR_DUMB_BASED, fresnel);
  const float mt = std::max(std::min(metalness, 1.0f), 0.0f);
  m.diffuse[0] = diffuse.x * (1.0f - mt); m.diffuse[1] = diffuse.y * (1.0f - mt); m.diffuse[2] = diffuse.z * (1.0f - mt);
  m.p[0] = specular;
  m.p[1] = mt;
  return m;
}
// Presets, DumbBasedShader.cuh:14-17.
inline dr_material roughGold() { return dumbBasedShader({1.022f, 0.782f, 0.344f}, 16.0f, {1.0f, 1.0f, 1.0f}, 1.0f); }
inline dr_material glossyGold() { return dumbBasedShader({1.022f, 0.782f, 0.344f}, 128.0f, {1.0f, 1.0f, 1.0f}, 1.0f); }
inline dr_material glossyFinish() { return dumbBasedShader({0.175f, 0.15f, 0.125f}, 4096.0f, {0.32f, 0.32f, 0.32f}, 0.25f); }
inline dr_material matteFinish() { return dumbBasedShader({0.175f, 0.15f, 0.125f}, 8.0f, {0.24f, 0.24f, 0.24f}, 0.05f); }
// diffuse and gloss divided by their sum when it exceeds 1 (SimpleStochasticShader.impl.cuh:4-11,
// DefaultShader.impl.cuh:5-13).
inline dr_material diffuseGloss(int32_t kind, Vector3 color, float diffuse, float smoothness, float shininess) {
  dr_material m = material(kind, color);
  const float total = diffuse + smoothness;
  const float mul = (total > 1.0f) ? (1.0f / total) : 1.0f;
  m.p[0] = diffuse * mul;
  m.p[1] = smoothness * mul;
  m.p[2] = shininess;
  return m;
}
inline dr_material simpleStochasticShader(Vector3 color = {1.0f, 1.0f, 1.0f}, float diffuse = 0.05f,
                                          float smoothness = 0.95f, float shininess = 2.0f) {
  return diffuseGloss(DR_SIMPLE_STOCHASTIC, color, diffuse, smoothness, shininess);
}
inline dr_material fresnelShader(Vector3 color = {1.0f, 1.0f, 1.0f}) { return material(DR_FRESNEL, color); }
inline dr_material defaultShader(Vector3 color = {1.0f, 1.0f, 1.0f}, float diffuse = 0.7f, float smoothness = 0.2f,
                                 float shine = 8.0f) {
  return diffuseGloss(DR_DEFAULT_SHADER, color, diffuse, smoothness, shine);
}
inline Vector3 normalized(Vector3 v) {   // Vector3::normalized
  const float inv = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  return {v.x * inv, v.y * inv, v.z * inv};
}
inline dr_light areaLight(Vector3 corner, Vector3 edgeU, Vector3 edgeV, Vector3 emission) {
  dr_light l{};
  l.kind = DR_AREA_LIGHT;
  const float v[12] = {corner.x, corner.y, corner.z, edgeU.x, edgeU.y, edgeU.z,
                       edgeV.x,  edgeV.y,  edgeV.z,  emission.x, emission.y, emission.z};
  for (int k = 0; k < 3; k++) {
    l.corner[k] = v[k]; l.edge_u[k] = v[3 + k]; l.edge_v[k] = v[6 + k]; l.emission[k] = v[9 + k];
  }
  return l;
}
// SimpleDirectionalLight (SimpleDirectionalLight.impl.cuh:7-11)
inline dr_light directionalLight(Vector3 color, Vector3 direction, float distance) {
  dr_light l{};
  l.kind = DR_DIRECTIONAL_LIGHT;
  const Vector3 d = normalized(direction);
  l.edge_u[0] = d.x; l.edge_u[1] = d.y; l.edge_u[2] = d.z;
  l.emission[0] = color.x; l.emission[1] = color.y; l.emission[2] = color.z;
  l.dist = distance;
  return l;
}
// SimpleSoftDirectionalLight (SimpleSoftDirectionalLight.impl.cuh:5-10)
inline dr_light softDirectionalLight(Vector3 color = {1.0f, 1.0f, 1.0f}, Vector3 direction = {1.0f, -1.0f, 1.0f},
                                     float distance = 1024.0f, float softness = 0.125f) {
  dr_light l = directionalLight(color, direction, distance);
  l.kind = DR_SOFT_DIRECTIONAL_LIGHT;
  l.radius = distance * softness;
  return l;
}
// SphericalLightEmitter (SphericalLightEmitter.impl.cuh:5-9)
inline dr_light sphericalLight(Vector3 color = {1.0f, 1.0f, 1.0f}, Vector3 position = {0.0f, 0.0f, 0.0f},
                               float radius = 1.0f) {
  dr_light l{};
  l.kind = DR_SPHERICAL_LIGHT;
  l.corner[0] = position.x; l.corner[1] = position.y; l.corner[2] = position.z;
  l.emission[0] = color.x; l.emission[1] = color.y; l.emission[2] = color.z;
  l.radius = radius;
  return l;
}
// Lights::Spotlight(shade, lum, pos, dir, innerAngle, outerAngle, falloffPower, emitSize,
// emitFromSphere, castShadows, samples) as its constructor stores it (Spotlight.impl.cuh:7-38).
inline dr_light spotlight(Vector3 color = {1.0f, 1.0f, 1.0f}, float luminosity = 1500.0f,
                          Vector3 position = {0.0f, 0.0f, 0.0f}, Vector3 direction = {0.0f, 0.0f, 1.0f},
                          float innerAngle = 16.0f, float outerAngle = 64.0f, float falloffPower = 1.0f,
                          float emitSize = 0.25f, bool emitFromSphere = false, bool castShadows = true, int samples = 1) {
  dr_light l{};
  l.kind = DR_SPOT_LIGHT;
  l.emission[0] = color.x * luminosity; l.emission[1] = color.y * luminosity; l.emission[2] = color.z * luminosity;
  l.corner[0] = position.x; l.corner[1] = position.y; l.corner[2] = position.z;
  const float inv = 1.0f / std::sqrt(direction.x * direction.x + direction.y * direction.y + direction.z * direction.z);
  l.edge_u[0] = direction.x * inv; l.edge_u[1] = direction.y * inv; l.edge_u[2] = direction.z * inv;
  if (emitSize < (32 * 0.000005f)) emitSize = (32 * 0.000005f);
  if (innerAngle < 0.0f) innerAngle = (-innerAngle);
  if (innerAngle > 180.0f) innerAngle = 180.0f;
  if (outerAngle < 0.0f) outerAngle = (-innerAngle);
  if (outerAngle > 180.0f) outerAngle = 180.0f;
  if (outerAngle < innerAngle) outerAngle = innerAngle;
  l.edge_v[0] = std::cos(innerAngle * 3.14159265359f / 180.0f);
  l.edge_v[1] = std::cos(outerAngle * 3.14159265359f / 180.0f);
  l.edge_v[2] = falloffPower;
  l.radius = emitSize;
  l.reserved = ((samples > 0) ? ((samples > 16) ? 16 : samples) : 0) | (castShadows ? 32 : 0) | (emitFromSphere ? 64 : 0);
  return l;
}
}  // namespace Plugins

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
  // DumbRenderer::BoxingMode (DumbRenderer.cuh:12-18)
  enum BoxingMode {
    BOXING_MODE_HEIGHT_BASED = DR_BOXING_HEIGHT,
    BOXING_MODE_WIDTH_BASED = DR_BOXING_WIDTH,
    BOXING_MODE_MIN_BASED = DR_BOXING_MIN,
    BOXING_MODE_MAX_BASED = DR_BOXING_MAX
  };
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
    // the reference estimator as soon as any of the reference's own plugins is used
    plugins_ = DR_PLUGINS_BENCHMARK;
    for (const dr_material &m : materials)
      if (m.kind >= DR_DUMB_BASED) plugins_ = DR_PLUGINS_REFERENCE;
    for (const dr_light &l : lights)
      if (l.kind != DR_AREA_LIGHT) plugins_ = DR_PLUGINS_REFERENCE;
  }
  ~DumbRenderer() {
    (void)hipFree(mats_);
    (void)hipFree(lights_);
    (void)hipFree(maps_);
    (void)hipFree(textures_);
    (void)hipFree(texels_);
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
  void setLens(int lens) { lens_ = lens; }   // DR_STOCHASTIC_LENS, DR_PERSPECTIVE_LENS or DR_SPHERICAL_SEGMENT_LENS
  // Lenses::SphericalSegmentLense parameters (angles, focal distance, sensor size, filter)
  // and the camera Transform's point-map rows; without rows, the camera's right, up, forward.
  void setLensParams(const float lensP[8], const float *camRows = nullptr) {
    for (int k = 0; k < 8; k++) lensP_[k] = lensP[k];
    hasRows_ = camRows != nullptr;
    if (camRows)
      for (int k = 0; k < 9; k++) camRows_[k] = camRows[k];
  }
  // Texture tables (ColoredTexture maps of the materials that have them, the textures and
  // their RGBA fp32 texels), copied to the device.
  void setTextures(const std::vector<dr_material_maps> &maps, const std::vector<dr_texture> &textures,
                   const std::vector<float> &texels) {
    upload(maps_, maps);
    upload(textures_, textures);
    upload(texels_, texels);
    nmaps_ = (int)maps.size();
    ntextures_ = (int)textures.size();
  }
  void setBoxingMode(BoxingMode mode) { boxing_ = mode; }
  dr_scene deviceScene() const { return scene_; }
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
    for (int k = 0; k < 9; k++) p.cam_rows[k] = hasRows_ ? camRows_[k] : cam[3 + k];
    for (int k = 0; k < 8; k++) p.lens_p[k] = lensP_[k];
    p.maps = maps_;
    p.textures = textures_;
    p.texels = texels_;
    p.nmaps = nmaps_;
    p.ntextures = ntextures_;
    p.materials = mats_;
    p.lights = lights_;
    p.nmat = nmat_;
    p.nlights = nlights_;
    p.lens = lens_;
    p.plugins = plugins_;
    p.boxing = boxing_;
    return p;
  }
  // BufferedRenderer / Renderer iteration interface (Renderer.cu:136-156, BufferedRenderer.cu:7):
  // iterate() renders iteration k + 1 into the frame buffer (blending 1/(k + 1)) and waits for it;
  // the resolution follows the frame buffer.
  void setFrameBuffer(BlockBuffer *frame) {
    if (frame != frame_) {
      frame_ = frame;
      resetIterations();
    }
  }
  BlockBuffer *getFrameBuffer() const { return frame_; }
  int iteration() const { return iteration_; }
  void resetIterations() { iteration_ = 0; }
  bool iterate() {
    if (!frame_) return false;
    int w, h;
    frame_->getSize(&w, &h);
    const dr_render_params p = params(w, h, iteration_ + 1);
    if (dr_render_blocks(scene_, &p, frame_->deviceData(), 0, frame_->getBlockCount(), nullptr) != DR_OK) return false;
    if (hipStreamSynchronize(nullptr) != hipSuccess) return false;
    iteration_++;
    return true;
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
  template <typename T, typename D>
  static void upload(D *&dev, const std::vector<T> &v) {
    (void)hipFree(dev);
    dev = nullptr;
    if (v.empty()) return;
    hip_check(hipMalloc((void **)&dev, sizeof(T) * v.size()), "hipMalloc");
    hip_check(hipMemcpy((void *)dev, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "copy");
  }
  dr_scene scene_;
  Camera camera_;
  dr_material *mats_ = nullptr;
  dr_light *lights_ = nullptr;
  dr_material_maps *maps_ = nullptr;
  dr_texture *textures_ = nullptr;
  float *texels_ = nullptr;
  int nmaps_ = 0, ntextures_ = 0;
  float lensP_[8] = {0, 0, 0, 0, 0, 0, 0, 0}, camRows_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool hasRows_ = false;
  int nmat_, nlights_;
  int bounces_ = 2, fx_ = 1, fy_ = 1, rrStart_ = 1 << 30;
  float rrSurvive_ = 1.0f;
  bool ignoreBackfaces_ = true;
  uint64_t seed_ = 0x5EED;
  int lens_ = DR_STOCHASTIC_LENS, plugins_ = DR_PLUGINS_BENCHMARK;
  int boxing_ = DR_BOXING_HEIGHT;
  BlockBuffer *frame_ = nullptr;
  int iteration_ = 0;
};

// Many renders in one persistent launch captured in a hipGraph (dr_batch_*), in place of
// BlockRenderer's host-driven block batches (BlockRenderer.cu:74-97).  Requests of both
// plugin sets and every lens may be mixed; each frame ends as renderBlocksGPU would leave it.
class RenderBatch {
 public:
  struct Request {
    DumbRenderer *renderer;
    BlockBuffer *frame;
    int iteration;
    uint32_t request;        // Philox counter word 3
  };
  // tuning == nullptr runs the autotuner
  explicit RenderBatch(const std::vector<Request> &requests, const dr_batch_tuning *tuning = nullptr) {
    std::vector<dr_request> r(requests.size());
    for (size_t i = 0; i < requests.size(); i++) {
      int w, h;
      requests[i].frame->getSize(&w, &h);
      r[i].scene = requests[i].renderer->deviceScene();
      r[i].params = requests[i].renderer->params(w, h, requests[i].iteration, requests[i].request);
      r[i].frame = requests[i].frame->deviceData();
    }
    check(dr_batch_create(r.data(), (int)r.size(), tuning, &batch_), "dr_batch_create");
  }
  ~RenderBatch() { dr_batch_destroy(batch_); }
  RenderBatch(const RenderBatch &) = delete;
  RenderBatch &operator=(const RenderBatch &) = delete;
  void run(hipStream_t stream = nullptr) { check(dr_batch_run(batch_, (void *)stream), "dr_batch_run"); }
  dr_batch_tuning tuning() const {
    dr_batch_tuning t{};
    check(dr_batch_get_tuning(batch_, &t), "dr_batch_get_tuning");
    return t;
  }

 private:
  dr_batch batch_ = nullptr;
};

// Images::saveBufferPNG (Images.cu:22-36): the frame as an 8-bit RGBA PNG, each channel
// 255*c clamped and truncated; returns the reference's error codes.
namespace Images {
enum Error {
  IMAGES_NO_ERROR = 0,
  IMAGES_ERROR_FRAME_BUFFER_INVALID = 1,
  IMAGES_ERROR_ALLOCATION_FAILED = 2,
  IMAGES_ERROR_EXTERNAL_LIBRARY_FAILED = 3
};
inline Error saveBufferPNG(BlockBuffer &frame, const std::string &filename) {
  int w, h;
  frame.getSize(&w, &h);
  if (w <= 0 || h <= 0) return IMAGES_ERROR_FRAME_BUFFER_INVALID;
  uint8_t *dev = nullptr;
  if (hipMalloc((void **)&dev, (size_t)w * h * 4) != hipSuccess) return IMAGES_ERROR_ALLOCATION_FAILED;
  std::vector<uint8_t> host((size_t)w * h * 4);
  const bool ok = dr_frame_to_rgba8(frame.deviceData(), w, h, dev, nullptr) == DR_OK &&
                  hipMemcpy(host.data(), dev, host.size(), hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(dev);
  if (!ok) return IMAGES_ERROR_EXTERNAL_LIBRARY_FAILED;
  return dr_png_write(filename.c_str(), host.data(), w, h) == DR_OK ? IMAGES_NO_ERROR
                                                                    : IMAGES_ERROR_EXTERNAL_LIBRARY_FAILED;
}
}  // namespace Images

// BufferedRenderProcess (BufferedRenderProcess.cu:16-229): a host thread that keeps
// iterating a renderer into the back buffer until the target iteration count (-1 = no
// limit), swapping front and back after each iteration when they differ, resizing the
// back buffer to the target resolution (which resets the iterations), and calling the
// iteration / render-completion / already-rendered / error callbacks.  No window: a
// negative target resolution keeps the buffer's size.
class BufferedRenderProcess {
 public:
  typedef void (*Callback)(void *);
  BufferedRenderProcess() = default;
  ~BufferedRenderProcess() { end(); }
  BufferedRenderProcess(const BufferedRenderProcess &) = delete;
  BufferedRenderProcess &operator=(const BufferedRenderProcess &) = delete;

  void start() {
    std::lock_guard<std::mutex> g(threadLock_);
    if (started_) return;
    {
      std::lock_guard<std::mutex> s(settingsLock_);
      kill_ = false;
      started_ = true;
    }
    startClock_ = clock();
    thread_ = std::thread(&BufferedRenderProcess::renderProcess, this);
  }
  void end() {
    std::lock_guard<std::mutex> g(threadLock_);
    if (!started_) return;
    {
      std::lock_guard<std::mutex> s(settingsLock_);
      kill_ = true;
    }
    thread_.join();
    renderClocks_ = renderTime();
    std::lock_guard<std::mutex> s(settingsLock_);
    started_ = false;
  }
  void lockSettings() { settingsLock_.lock(); }
  void unlockSettings() { settingsLock_.unlock(); }
  void setRenderer(DumbRenderer *r) { std::lock_guard<std::mutex> s(settingsLock_); renderer_ = r; }
  void setBuffer(BlockBuffer *b) { std::lock_guard<std::mutex> s(settingsLock_); front_ = back_ = b; }
  void setDoubleBuffers(BlockBuffer *a, BlockBuffer *b) { std::lock_guard<std::mutex> s(settingsLock_); front_ = a; back_ = b; }
  void setTargetIterations(int n) { std::lock_guard<std::mutex> s(settingsLock_); iterations_ = n; }
  void setInfinateTargetIterations() { setTargetIterations(-1); }
  void setTargetResolution(int w, int h) { std::lock_guard<std::mutex> s(settingsLock_); tw_ = w; th_ = h; }
  void setIterationCompletionCallback(Callback c, void *a) { std::lock_guard<std::mutex> s(settingsLock_); onIter_ = {c, a}; }
  void setRenderCompletionCallback(Callback c, void *a) { std::lock_guard<std::mutex> s(settingsLock_); onDone_ = {c, a}; }
  void setAlreadyRenderedCallback(Callback c, void *a) { std::lock_guard<std::mutex> s(settingsLock_); onAlready_ = {c, a}; }
  void setErrorOnIteration(Callback c, void *a) { std::lock_guard<std::mutex> s(settingsLock_); onError_ = {c, a}; }
  // Returns once the render thread has read the current settings.
  void synchSettings() {
    std::unique_lock<std::mutex> lock(settingsLock_);
    if (!started_) return;
    synchCond_.wait(lock);
  }
  long long renderTime() const { return renderClocks_ + (started_ ? (long long)(clock() - startClock_) : 0); }

 private:
  struct Cb {
    Callback f = nullptr;
    void *arg = nullptr;
    void operator()() const { if (f) f(arg); }
  };
  void renderProcess() {
    bool swap = false;
    while (true) {
      DumbRenderer *renderer;
      BlockBuffer *front, *back;
      int target, tw, th;
      Cb onIter, onDone, onAlready, onError;
      {
        std::lock_guard<std::mutex> s(settingsLock_);
        if (kill_) break;
        renderer = renderer_;
        front = swap ? back_ : front_;
        back = swap ? front_ : back_;
        target = iterations_;
        tw = tw_;
        th = th_;
        onIter = onIter_;
        onDone = onDone_;
        onAlready = onAlready_;
        onError = onError_;
        synchCond_.notify_all();
      }
      if (back != nullptr && tw > 0 && th > 0) {
        int w, h;
        back->getSize(&w, &h);
        if (w != tw || h != th) {
          back->setResolution(tw, th);
          if (renderer) {
            renderer->resetIterations();
            renderClocks_ = 0;
            startClock_ = clock();
          }
        }
      }
      if (renderer == nullptr) {
        std::this_thread::sleep_for(std::chrono::milliseconds(8));
        continue;
      }
      renderer->setFrameBuffer(back);
      if (renderer->iteration() != target) {
        if (!renderer->iterate()) {
          onError();
          continue;
        }
        if (back != front) {
          swap = !swap;
          renderer->resetIterations();
          renderClocks_ = 0;
          startClock_ = clock();
        }
        onIter();
        if (renderer->iteration() == target) onDone();
      } else {
        onAlready();
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
      }
    }
    synchCond_.notify_all();
  }

  std::mutex threadLock_, settingsLock_;
  std::condition_variable synchCond_;
  std::thread thread_;
  std::atomic<bool> started_{false};
  bool kill_ = false;
  DumbRenderer *renderer_ = nullptr;
  BlockBuffer *front_ = nullptr, *back_ = nullptr;
  int iterations_ = -1, tw_ = -1, th_ = -1;
  Cb onIter_, onDone_, onAlready_, onError_;
  std::atomic<long long> renderClocks_{0};
  std::atomic<clock_t> startClock_{0};
};

// DumbRenderContext (DumbRenderContext.cuh): a .dumb scene file loaded (dr_scene_file_*)
// into an Octree of Renderable<BakedTriFace> with its materials, lights, camera and
// renderer settings.  buildFromFile mirrors DumbRenderContext::buildFromFile
// (DumbRenderContext.cu:295-301): on failure it returns false and writes the error to
// errorStream when one is given.
class DumbRenderContext {
 public:
  template <typename Stream>
  bool buildFromFile(const std::string &filename, Stream *errorStream) {
    dr_scene_file f = nullptr;
    if (dr_scene_file_load(filename.c_str(), &f) != DR_OK) {
      if (errorStream) (*errorStream) << dr_last_error() << "\n";
      return false;
    }
    int ntris = 0, nmat = 0, nlights = 0, nt = 0, nm = 0;
    int64_t ntx = 0;
    dr_scene_file_info(f, &settings_, &ntris, &nmat, &nlights);
    std::vector<float> tris(9 * (size_t)ntris), nrm(9 * (size_t)ntris), uv(6 * (size_t)ntris);
    std::vector<int32_t> mat(ntris);
    materials_.assign(nmat, dr_material{});
    lights_.assign(nlights, dr_light{});
    dr_scene_file_geometry(f, tris.data(), nrm.data(), mat.data());
    if (ntris > 0) dr_scene_file_texcoords(f, uv.data());
    dr_scene_file_tables(f, materials_.data(), lights_.data());
    // texture and map tables for the renderers
    dr_scene_file_texture_info(f, &nt, &ntx, &nm);
    textures_.assign(nt, dr_texture{});
    texels_.assign(4 * (size_t)ntx, 0.0f);
    maps_.assign(nm, dr_material_maps{});
    dr_scene_file_textures(f, textures_.data(), texels_.data(), maps_.data());
    dr_scene_file_destroy(f);
    geometry_.reset();
    for (int i = 0; i < ntris; i++) {
      Renderable<BakedTriFace> r{};
      const float *t = &tris[9 * (size_t)i], *n = &nrm[9 * (size_t)i];
      r.object.vert = Triangle{{t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}};
      r.object.norm = Triangle{{n[0], n[1], n[2]}, {n[3], n[4], n[5]}, {n[6], n[7], n[8]}};
      r.materialId = mat[i];
      geometry_.push(r);
    }
    geometry_.build();
    normals_ = std::move(nrm);
    if (ntris > 0) {
      check(dr_scene_set_normals(geometry_.deviceScene(), normals_.data(), ntris), "dr_scene_set_normals");
      check(dr_scene_set_texcoords(geometry_.deviceScene(), uv.data(), ntris), "dr_scene_set_texcoords");
    }
    return true;
  }
  bool buildFromFile(const std::string &filename) { return buildFromFile<std::ostream>(filename, nullptr); }

  const Octree<> &geometry() const { return geometry_; }
  const std::vector<dr_material> &materials() const { return materials_; }
  const std::vector<dr_light> &lights() const { return lights_; }
  const dr_scene_file_settings &settings() const { return settings_; }
  Camera camera() const {
    const float *c = settings_.cam;
    return Camera{{c[0], c[1], c[2]}, {c[3], c[4], c[5]}, {c[6], c[7], c[8]}, {c[9], c[10], c[11]}, c[12]};
  }
  // A renderer with the file's settings (DumbRenderContext::runWindowRender,
  // DumbRenderContext.cu:869-874).
  DumbRenderer *newRenderer() const {
    auto *r = new DumbRenderer(geometry_, materials_, lights_, camera(), settings_.max_bounces, settings_.fsaa_x,
                               settings_.fsaa_y, settings_.ignore_backfaces != 0);
    r->setLens(settings_.lens);
    r->setLensParams(settings_.lens_p, settings_.cam_rows);
    r->setBoxingMode((DumbRenderer::BoxingMode)settings_.boxing);
    r->setTextures(maps_, textures_, texels_);
    return r;
  }

 private:
  Octree<> geometry_;
  std::vector<dr_material> materials_;
  std::vector<dr_light> lights_;
  std::vector<float> normals_;
  std::vector<dr_texture> textures_;
  std::vector<float> texels_;
  std::vector<dr_material_maps> maps_;
  dr_scene_file_settings settings_{};
};

}  // namespace dumbray
