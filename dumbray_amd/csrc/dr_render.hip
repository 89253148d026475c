// Path tracing: DumbRenderer::renderBlocksGPU and Renderer::backward on gfx950.
//
// The work of a render is a list of path jobs (pixel, sample) per 8x16 framebuffer
// block; idle lanes take the next jobs with a wave64 ballot + mbcnt prefix and one LDS
// atomic per wave, so lanes whose path ended early are regenerated instead of idling.
// Each lane's current ray (camera/bounce ray or light photon) is a Walk advanced by one
// node pop per loop iteration; leaf tests and the path updates of finished casts run
// once enough lanes of the wave need them, which keeps the wave's lanes in step.
// Finished sample colours are staged in LDS and summed per pixel in the reference's
// subpixel order before the pixel is written (set, or blended with 1/iteration).
// k_render_stream runs the jobs of many blocks as one stream per persistent workgroup;
// render_block (k_render, k_batch) runs one block per workgroup at a time.
#include "dr_common.h"
#include "dr_plugins.h"
#include "dr_shade.h"

#include <algorithm>
#include <new>
#include <vector>

namespace dr {

constexpr int kRB = 256;               // lanes per workgroup
constexpr int kRoundSamples = 16;      // samples per pixel staged in LDS per round
constexpr int kBlockPixels = 128;      // BlockBasedFrameBuffer<8,16>
constexpr int kLeafMin = 16;           // pending leaves that trigger the (wave-shared) leaf tests
constexpr int kShadeMin = 32;          // finished casts that trigger the adjoint's path updates
// Idle lanes a wave collects before it starts new path jobs (1, 8, 16, 32 measured
// 34.9 / 35.0 / 35.3 / 38.4 ms at 1080p 16 spp, 0.84 / 0.77 / 0.80 / 0.90 ms at 256x256
// 4 spp, 15.7 / 15.3 / 15.6 / 16.3 ms for the example scene at 1 spp, one block per
// workgroup; 4, 8, 16 gave 29.8 / 29.8 / 30.4 ms at 1080p on the streamed kernel).  Used by the
// adjoint and as the batch autotuner's starting point; the render kernels take the estimator's
// kRefillMin / kStreamRefillMin.
constexpr int kRenderRefillMin = 8;
// Waves per SIMD the render kernels are compiled for: the register cap this sets
// (168 VGPRs) costs a few spills but beats 2 waves per SIMD clearly.
constexpr int kRenderWaves = 3;

// Per-lane path state.
struct Path {
  int job = -1;
  uint32_t pix = 0, samp = 0;
  int layer = 0, phase = 0, light = 0;  // phase 0: geometry ray, 1: light photon
  V3 o, d;
  int restrict_tri = -1;
  float beta[3], L[3];
  int h = -1, kind = 0, mid = 0;
  V3 p, nu;
  bool has_sample = false;
  V3 sdir;
  float sweight = 0.0f;
  float contrib[3];                     // photon contribution if lit
  float gl_k = 0.0f;                    // (adjoint) cos_i/pi * cos_l*A/d^2 of the pending photon
};

// A value that is the same in every lane (a function of the kernel's parameters), moved to
// a scalar register.  Call with every lane that holds it agreeing on it.
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ UDiv uni(UDiv D) { return {uni(D.d), uni(D.m)}; }
__device__ __forceinline__ UDiv udiv_uniform(uint32_t d) { return uni(make_udiv(d)); }
// frame_k(P) formed in the kernel and moved to scalar registers (the kernels otherwise take
// the host's frame_k as an argument).
__device__ __forceinline__ FrameK frame_k_uniform(const dr_render_params &P) {
  FrameK K = frame_k(P);
  K.nbw = uni(K.nbw);
  K.fy = uni(K.fy);
  K.hw = uni(K.hw);
  K.hh = uni(K.hh);
  K.ps = uni(K.ps);
  K.pdx = uni(K.pdx);
  K.pdy = uni(K.pdy);
  K.jx = uni(K.jx);
  K.ox = uni(K.ox);
  K.jy = uni(K.jy);
  K.oy = uni(K.oy);
  for (int k = 0; k < 4; k++) K.zc[k] = uni(K.zc[k]);
  K.amount = uni(K.amount);
  K.keep = uni(K.keep);
  K.inv_rr = uni(K.inv_rr);
  return K;
}

// Common job start: camera ray of sample `s` of local pixel `lp`.  Returns false
// (and leaves the lane idle) for pixels outside the image.
__device__ __forceinline__ bool start_path(const dr_render_params &P, const FrameK &K, int block, int lp, int s,
                                           uint32_t samp0, Path &st) {
  int x, y;
  if (!block_pixel(K, P.width, P.height, block, lp, x, y)) return false;
  st.pix = (uint32_t)y * (uint32_t)P.width + (uint32_t)x;
  st.samp = samp0 + (uint32_t)s;
  const int i = (int)udiv((uint32_t)s, K.fy), j = s - i * P.fsaa_y;
  float u[4];
  rng4(P, st.pix, st.samp, 0, 0, u);
  camera_ray(P, K, x, y, i, j, u[0], u[1], u[2], u[3], st.o, st.d);
  st.layer = 0;
  st.phase = 0;
  st.restrict_tri = -1;
  for (int c = 0; c < 3; c++) { st.beta[c] = 1.0f; st.L[c] = 0.0f; }
  return true;
}

// Sets up the next light photon of the current vertex (phase 1), or returns false.
__device__ __forceinline__ bool next_photon(const dr_render_params &P, Path &st, float *cinfo = nullptr) {
  for (int l = st.light + 1; l < P.nlights; l++) {
    const dr_light Lt = P.lights[l];
    V3 q, w;
    float cos_i, gl;
    if (!light_photon(P, Lt, l, st.pix, st.samp, st.layer, st.p, st.nu, q, w, cos_i, gl)) continue;
    const dr_material M = P.materials[st.mid];
    const float k = cos_i * kInvPi;
    for (int c = 0; c < 3; c++) st.contrib[c] = st.beta[c] * ((M.albedo[c] * k) * (Lt.emission[c] * gl));
    st.gl_k = k * gl;
    st.light = l;
    st.phase = 1;
    st.o = q;
    st.d = w;
    st.restrict_tri = -1;
    return true;
  }
  return false;
}

// Vertex setup after a geometry hit of the ray with direction `dir`.
__device__ __forceinline__ void setup_vertex(const dr_render_params &P, const FrameK &K, const float4 *__restrict__ shade,
                                             const Hit &hh, V3 dir, Path &st) {
  st.h = hh.id;
  st.p = hh.p;
  const float4 sh = shade[hh.id];
  st.mid = __float_as_int(sh.w);
  st.kind = P.materials[st.mid].kind;
  const V3 du = normalized(dir);
  st.nu = facing_normal(v3(sh.x, sh.y, sh.z), du);
  st.has_sample = next_sample(P, K, st.kind, st.pix, st.samp, st.layer, st.nu, du, st.sdir, st.sweight);
  st.light = -1;
}

// Reflection factor A of the current vertex for its sampled direction.
__device__ __forceinline__ void reflect_factor(const dr_render_params &P, const Path &st, float A[3]) {
  const dr_material M = P.materials[st.mid];
  if (st.kind == DR_METAL) {
    const float f5 = schlick5(st.nu, st.sdir);
    for (int c = 0; c < 3; c++) A[c] = M.albedo[c] + (1.0f - M.albedo[c]) * f5;
  } else {
    for (int c = 0; c < 3; c++) A[c] = M.albedo[c];
  }
}

// Shading tables of one scene.
struct Tables {
  const float4 *shade;   // per triangle: (n, material id bits)
  const float *faces;    // per triangle: kFaceWords floats (vertices, vertex normals)
};

// Path updates of the benchmark plugins (DR_PLUGINS_BENCHMARK).
struct BenchPath {
  static constexpr bool kFastSlab = true;
  // k_render_stream: a job outside the image is written at once, and the job's block is read
  // back from the block ring when its pixel is written (29.5 ms at 1080p 16 spp, against
  // 30.8 ms with the other choices, which suit RefPath)
  static constexpr bool kStreamBlackInUpdate = false, kStreamBlockInRegister = false;
  // Finished casts that trigger the path updates and idle lanes that trigger new path jobs, in
  // render_block and in k_render_stream (1080p 16 spp, split 8 / depth 24: 27.2 ms at 32 / 8,
  // 26.4-26.5 ms at 24 or 20 / 4 in the streamed kernel; 0.67 ms at 32 / 8 against 0.70 ms at
  // 24 / 4 for the 256x256 frame in render_block)
  static constexpr int kShadeMin = 32, kRefillMin = 8, kStreamShadeMin = 24, kStreamRefillMin = 4;
  static constexpr bool kStreamFrameInKernel = false;
  Path st;
  __device__ __forceinline__ bool start(const dr_render_params &P, const FrameK &K, int block, int lp, int s,
                                        uint32_t samp0) {
    return start_path(P, K, block, lp, s, samp0, st);
  }
  __device__ __forceinline__ V3 ray_o() const { return st.o; }
  __device__ __forceinline__ V3 ray_d() const { return st.d; }
  __device__ __forceinline__ int ray_excl() const { return st.restrict_tri; }
  __device__ __forceinline__ const float *color() const { return st.L; }
  __device__ __forceinline__ bool photon() const { return st.phase == 1; }
  // Next bounce after the last photon of a vertex; false when the path ends.
  __device__ __forceinline__ bool advance(const dr_render_params &P) {
    if (!st.has_sample) return false;
    float A[3];
    reflect_factor(P, st, A);
    for (int c = 0; c < 3; c++) st.beta[c] = st.beta[c] * (A[c] * st.sweight);
    st.o = st.p;
    st.d = st.sdir;
    st.restrict_tri = st.h;
    st.layer++;
    st.phase = 0;
    return true;
  }
  // After the lane's cast ended.  A geometry hit sets up the vertex, then photons
  // (Lambert) or the next bounce; a photon adds its contribution if it reached the
  // vertex, then the next photon.  Returns true when the path is done.
  __device__ __forceinline__ bool shade(const dr_render_params &P, const FrameK &K, const Tables &tb, const Walk &w) {
    bool photons = false;
    if (st.phase == 0) {
      if (!w.hit) return true;
      setup_vertex(P, K, tb.shade, w.h, w.ray.d, st);
      photons = st.kind == DR_LAMBERT;
    } else {
      const V3 e = w.h.p - st.p;
      const bool lit = w.hit && (w.h.id == st.h || (e.x * e.x + e.y * e.y + e.z * e.z) <= (8.0f * kVecEps));
      if (lit) for (int c = 0; c < 3; c++) st.L[c] += st.contrib[c];
      photons = true;
    }
    if (photons && next_photon(P, st)) return false;
    return !advance(P);
  }
};

// Path updates of the reference estimator (DR_PLUGINS_REFERENCE), DumbRenderer.cu:256-443:
// every layer casts each light's photons (one, or a spotlight's samples); a photon that first reaches another surface
// continues through it with that surface's getReflectedColor (observer = the photon
// direction) as its colour when that colour is not black (:396-413).  Colours are
// carried forward: L += beta * (getReflectedColor of each direct photon), and the
// child layer's throughput is beta * getReflectedColor(indirect photon) * sampleWeight.
//
// Only what the shadow casts need stays live across them: the vertex (triangle, point,
// incoming direction, material) and the photon.  The shading record and the indirect
// sample are computed again where they are used; both are functions of the vertex and
// of counter-based random numbers, so the order of evaluation does not change them.
struct RefPath {
  // waves holding an axis-parallel photon (a directional light along an axis) take the
  // NaN-handling form; the example scene's soft lights do not (15.0 -> 14.1 ms at 1080p)
  static constexpr bool kFastSlab = true;
  // k_render_stream: a job outside the image goes through the path-update phase, and the
  // job's block stays in a register (example scene 8.5 ms, against 9.4-9.9 ms otherwise:
  // fewer spilled registers)
  static constexpr bool kStreamBlackInUpdate = true, kStreamBlockInRegister = true;
  // the same thresholds (example scene, streamed: 8.12 ms at 32 / 8, 8.01 at 44 / 8, 7.94 at
  // 44 / 4, 9.37 at 64 / 8)
  static constexpr int kShadeMin = 32, kRefillMin = 8, kStreamShadeMin = 44, kStreamRefillMin = 4;
  static constexpr bool kStreamFrameInKernel = true;
  uint32_t pix, samp;
  int layer, phase, light, lk, excl, trans;   // lk: photons of the current light still to send
  V3 o, d;                          // the next ray to cast
  float beta[3], L[3];
  int h, mid;                       // current vertex: triangle, material
  V3 hp, din;                       //   hit point, incoming ray direction
  V3 pd;                            // current photon: direction, colour, unit direction?
  float pc[3];
  bool unit;

  __device__ __forceinline__ bool start(const dr_render_params &P, const FrameK &K, int block, int lp, int s,
                                        uint32_t samp0) {
    int x, y;
    if (!block_pixel(K, P.width, P.height, block, lp, x, y)) return false;
    pix = (uint32_t)y * (uint32_t)P.width + (uint32_t)x;
    samp = samp0 + (uint32_t)s;
    const int i = (int)udiv((uint32_t)s, K.fy), j = s - i * P.fsaa_y;
    float u[4];
    rng4(P, pix, samp, 0, 0, u);
    camera_ray(P, K, x, y, i, j, u[0], u[1], u[2], u[3], o, d);
    layer = 0;
    phase = 0;
    excl = -1;
    for (int c = 0; c < 3; c++) { beta[c] = 1.0f; L[c] = 0.0f; }
    return true;
  }
  __device__ __forceinline__ V3 ray_o() const { return o; }
  __device__ __forceinline__ V3 ray_d() const { return d; }
  __device__ __forceinline__ int ray_excl() const { return excl; }
  __device__ __forceinline__ const float *color() const { return L; }
  __device__ __forceinline__ bool photon() const { return phase == 1; }

  // The vertex's shading record (reference shaders; only the point for the others).
  __device__ __forceinline__ RefHit hit_record(const dr_render_params &P, const Tables &tb,
                                               const dr_material &M) const {
    RefHit H;
    H.p = hp;
    if (is_reference_kind(M.kind)) {
      const TexCtx T = tex_ctx(P);
      H = ref_hit(tb.faces, h, hp, T, material_maps(T, M));
    }
    return H;
  }
  __device__ __forceinline__ V3 facing(const Tables &tb) const {
    const float4 sh = tb.shade[h];
    return facing_normal(v3(sh.x, sh.y, sh.z), normalized(din));
  }
  // Direct light of the current photon at the vertex.
  __device__ __forceinline__ void add_direct(const dr_render_params &P, const Tables &tb) {
    const dr_material M = P.materials[mid];
    if (M.kind == DR_LAMBERT) {
      const V3 wn = unit ? pd : normalized(pd);
      const float cos_i = -dot(facing(tb), wn);
      if (cos_i > 0.0f) {
        const float k = cos_i * kInvPi;
        for (int c = 0; c < 3; c++) L[c] += beta[c] * ((M.albedo[c] * k) * pc[c]);
      }
    } else {
      const RefHit H = hit_record(P, tb, M);
      float f[3];
      reflect_ref(M, H, pd, pc, neg(din), true, 0, tex_ctx(P), f);
      for (int c = 0; c < 3; c++) L[c] += beta[c] * f[c];
    }
  }
  // The indirect sample of the vertex and the child layer's ray and throughput; false when
  // the path ends here.
  __device__ __forceinline__ bool next_layer(const dr_render_params &P, const FrameK &K, const Tables &tb) {
    if (layer >= P.max_bounces) return false;
    const dr_material M = P.materials[mid];
    float u[4];
    rng4(P, pix, samp, (uint32_t)layer, 1, u);
    float sw = 1.0f;
    if (layer >= P.rr_start) {
      if (u[2] >= P.rr_survive) return false;
      sw = K.inv_rr;
    }
    const RefHit H = hit_record(P, tb, M);
    V3 sdir;
    int stype = 0;
    float F[3];
    if (M.kind == DR_LAMBERT) {
      sdir = cosine_dir(facing(tb), u[0], u[1]);
      for (int c = 0; c < 3; c++) F[c] = M.albedo[c];
    } else if (M.kind == DR_MIRROR || M.kind == DR_METAL) {
      const V3 nu = facing(tb);
      sdir = reflect_unit(normalized(din), nu);
      if (M.kind == DR_METAL) {
        const float f5 = schlick5(nu, sdir);
        for (int c = 0; c < 3; c++) F[c] = M.albedo[c] + (1.0f - M.albedo[c]) * f5;
      } else {
        for (int c = 0; c < 3; c++) F[c] = M.albedo[c];
      }
    } else {
      if (!sample_ref(P, M, H, din, pix, samp, layer, u, sdir, stype)) return false;
      const float white[3] = {1.0f, 1.0f, 1.0f};
      reflect_ref(M, H, neg(sdir), white, neg(din), false, stype, tex_ctx(P), F);
    }
    for (int c = 0; c < 3; c++) beta[c] = beta[c] * (F[c] * sw);
    o = hp;
    d = sdir;
    excl = h;
    layer++;
    phase = 0;
    return true;
  }
  // After the lane's cast ended; true when the path is done.
  __device__ __forceinline__ bool shade(const dr_render_params &P, const FrameK &K, const Tables &tb, const Walk &w) {
    if (phase == 0) {
      if (!w.hit) return true;                        // the layer ray missed (:418-421)
      h = w.h.id;
      hp = w.h.p;
      din = w.ray.d;
      mid = __float_as_int(tb.shade[h].w);
      const int kind = P.materials[mid].kind;
      light = (kind == DR_MIRROR || kind == DR_METAL) ? P.nlights : -1;   // delta BSDFs take no direct light
      lk = 0;
    } else if (w.hit) {
      const V3 e = w.h.p - hp;
      if (w.h.id == h || (e.x * e.x + e.y * e.y + e.z * e.z) <= (8.0f * kVecEps)) {
        add_direct(P, tb);
      } else {
        const int hb = w.h.id;
        const dr_material Mb = P.materials[__float_as_int(tb.shade[hb].w)];
        if (is_reference_kind(Mb.kind) && trans < kMaxTransmit) {
          const TexCtx T = tex_ctx(P);
          const RefHit B = ref_hit(tb.faces, hb, w.h.p, T, material_maps(T, Mb));
          float tc[3];
          reflect_ref(Mb, B, pd, pc, pd, true, 0, T, tc);
          if (((tc[0] + tc[1]) + tc[2]) > kVecEps) {
            for (int c = 0; c < 3; c++) pc[c] = tc[c];
            o = w.h.p + normalized(pd) * (64.0f * kVecEps);
            d = pd;
            excl = hb;
            trans++;
            return false;
          }
        }
      }
    }
    // next photon: the current light's remaining samples (last first, DumbRenderer.cu:290-293),
    // then the following lights; a light that casts no shadows illuminates at once (:306-315)
    while (true) {
      if (lk == 0) {
        if (++light >= P.nlights) break;
        lk = light_samples(P.lights[light]);
        if (lk == 0) continue;
      }
      lk--;
      V3 q, wv;
      bool u1;
      float c3[3];
      if (!photon_ref(P, P.lights[light], light, lk, pix, samp, layer, hp, q, wv, c3, u1)) continue;
      pd = wv;
      unit = u1;
      for (int c = 0; c < 3; c++) pc[c] = c3[c];
      if (!light_casts_shadows(P.lights[light])) {
        add_direct(P, tb);
        continue;
      }
      o = q;
      d = wv;
      excl = -1;
      trans = 0;
      phase = 1;
      return false;
    }
    return !next_layer(P, K, tb);
  }
};

// ------------------------------------------------------------------ forward
// LDS used by one block render.
struct BlockLds {
  uint32_t *stk;        // this lane's first stack word ([word][lane])
  float *samp_col;      // kBlockPixels * kRoundSamples * 3 sample colours
  float *acc;           // kBlockPixels * 3 pixel sums
  int *next_job;
  unsigned long long *best;   // this wave's leaf-test slots (walk_leaf_wave)
  int *ins;             // this wave's inside-test counts (measurement launches only)
};

// Work counters of a render (measurement launches only), summed over the launch.
enum : int {
  kRcSegments = 0,      // casts: camera and bounce rays and light photons
  kRcPhotons,           //   of which light photons (shadow rays)
  kRcExpand,            // group expansions (8 child box tests)
  kRcLeaves,            // leaves tested
  kRcPlane,             // triangle plane tests
  kRcInside,            // triangle inside tests
  kRcPaths,             // path jobs (pixel samples) started
  kRcLive,              // children of expanded groups with a non-empty leaf below (octant tests)
  kRcSlab,              // children whose slab test ran
  kRcCount
};
struct RenderCounters {
  CastCounters cast = {0, 0, 0, 0, 0, 0};
  int segments = 0, photons = 0, paths = 0;
};
// Adds one lane's counters to the launch's totals (every lane of the wave calls it).
__device__ __forceinline__ void flush_counters(const RenderCounters &c, unsigned long long *out) {
  const int v[kRcCount] = {c.segments, c.photons,     c.cast.expand, c.cast.leaves, c.cast.plane,
                           c.cast.inside, c.paths, c.cast.live, c.cast.slab};
#pragma unroll
  for (int k = 0; k < kRcCount; k++) {
    unsigned long long x = (unsigned long long)v[k];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    if (__lane_id() == 0) atomicAdd(&out[k], x);
  }
}

// Framebuffer pixel `lp` of block `block`: the colour of this iteration, set (iteration 1)
// or blended with weight 1/iteration (DumbRenderer.cu:64, BlockBasedFrameBuffer.impl.cuh:125-156).
__device__ __forceinline__ void store_pixel(const dr_render_params &P, const FrameK &K, float *__restrict__ frame,
                                            int block, int lp, const float a[3]) {
  int x, y;
  if (!block_pixel(K, P.width, P.height, block, lp, x, y)) return;
  float4 *f = (float4 *)frame + (size_t)kBlockPixels * block + lp;
  const float4 col = make_float4(a[0], a[1], a[2], 1.0f);
  const float amount = K.amount;
  if (amount >= 1.0f) *f = col;
  else {
    const float4 old = *f;
    const float ia = K.keep;
    *f = make_float4((col.x * amount) + (old.x * ia), (col.y * amount) + (old.y * ia),
                     (col.z * amount) + (old.z * ia), (col.w * amount) + (old.w * ia));
  }
}

// Renders framebuffer block `block` of one request (all lanes of the workgroup call it),
// in rounds of kRoundSamples samples per pixel.  K = frame_k(P).
// refill_min: an idle lane takes a new job once at least this many lanes of its wave
// are idle (or the whole wave is).
template <class PathT, bool kStats = false>
__device__ __forceinline__ void render_block(const DevTree &T, const Tables &tb, const dr_render_params &P,
                                             const FrameK &K, float *__restrict__ frame, int block,
                                             const BlockLds &lds, int refill_min, RenderCounters *rc = nullptr) {
  CastCounters *cc = kStats ? &rc->cast : nullptr;
  const int spp = P.fsaa_x * P.fsaa_y;
  const uint32_t samp0 = (uint32_t)(P.iteration > 1 ? P.iteration - 1 : 0) * (uint32_t)spp;
  const float spw = uni(1.0f / (((float)P.fsaa_x) * ((float)P.fsaa_y)));
  const bool clip = P.cull_backfaces != 0;
  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  float *samp_col = lds.samp_col;
  for (int t = threadIdx.x; t < kBlockPixels * 3; t += kRB) lds.acc[t] = 0.0f;

  for (int s_base = 0; s_base < spp; s_base += kRoundSamples) {
    const int ns = min(kRoundSamples, spp - s_base);
    const UDiv Dns = udiv_uniform((uint32_t)ns);
    const int njobs = kBlockPixels * ns;   // job j: pixel j / ns, sample s_base + j % ns
    if (threadIdx.x == 0) *lds.next_job = 0;
    __syncthreads();
    PathT ps;
    int job = -1;
    Walk w;
    w.state = kWalkDone;
    bool exhausted = false;
    auto finish = [&](const float *L) {
      // getPixelColor: the layer colour (SimpleStochasticLense, DefaultPerspectiveLense) or the
      // colour times the lens filter (SphericalSegmentLense.impl.cuh:31-34)
      if (P.lens == DR_SPHERICAL_SEGMENT_LENS)
        for (int c = 0; c < 3; c++) samp_col[3 * job + c] = L[c] * P.lens_p[4 + c];
      else
        for (int c = 0; c < 3; c++) samp_col[3 * job + c] = L[c];
      job = -1;
    };
    for (;;) {
      const bool want = (job < 0) && !exhausted;
      const uint64_t wmask = __ballot(want);
      const uint64_t busy = __ballot(job >= 0);
      if (wmask && (__popcll(wmask) >= refill_min || busy == 0)) {
        const int leader = __ffsll((unsigned long long)wmask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(lds.next_job, __popcll(wmask));
        base = __shfl(base, leader);
        if (want) {
          const int slot = base + __popcll(wmask & below);
          if (slot < njobs) {
            job = slot;
            const int rp = (int)udiv((uint32_t)slot, Dns), s = s_base + (slot - rp * ns);
            if (!ps.start(P, K, block, rp, s, samp0)) {
              const float zero[3] = {0.0f, 0.0f, 0.0f};
              finish(zero);
            } else {
              walk_begin(T, w, ps.ray_o(), ps.ray_d(), ps.ray_excl(), lds.stk, kRB, nullptr);
              if (kStats) { rc->paths++; rc->segments++; }
            }
          } else {
            exhausted = true;
          }
        }
      }
      if (__ballot(job >= 0) == 0) {
        if (__ballot(!exhausted) == 0) break;
        continue;
      }
      // ---- one traversal step of every ray in flight
      if (job >= 0 && w.state == kWalkTrace) walk_step<PathT::kFastSlab>(T, w, lds.stk, kRB, cc);
      // ---- leaf tests, batched
      const bool pend = job >= 0 && w.state == kWalkLeaf;
      const uint64_t pm = __ballot(pend);
      if (pm && (__popcll(pm) >= kLeafMin || __ballot(job >= 0 && w.state == kWalkTrace) == 0))
        walk_leaf_wave(T, w, pend, clip, lds.best, kStats ? lds.ins : nullptr, cc);
      // ---- path updates of the finished casts, batched
      const uint64_t dm = __ballot(job >= 0 && w.state == kWalkDone);
      if (dm && (__popcll(dm) >= PathT::kShadeMin || __ballot(job >= 0 && w.state != kWalkDone) == 0)) {
        if (job >= 0 && w.state == kWalkDone) {
          if (ps.shade(P, K, tb, w)) {
            finish(ps.color());
          } else {
            walk_begin(T, w, ps.ray_o(), ps.ray_d(), ps.ray_excl(), lds.stk, kRB, nullptr);
            if (kStats) { rc->segments++; rc->photons += ps.photon() ? 1 : 0; }
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < kBlockPixels) {
      for (int s = 0; s < ns; s++)
        for (int c = 0; c < 3; c++) lds.acc[3 * threadIdx.x + c] += samp_col[3 * (threadIdx.x * ns + s) + c] * spw;
    }
    __syncthreads();
  }
  if (threadIdx.x < kBlockPixels) store_pixel(P, K, frame, block, threadIdx.x, lds.acc + 3 * threadIdx.x);
  __syncthreads();
}

// One framebuffer block per workgroup (sample counts above the streamed kernel's limit).
template <class PathT, bool kStats = false>
__global__ __launch_bounds__(kRB, kRenderWaves) void k_render(DevTree T, Tables tb, dr_render_params P, FrameK K,
                                                               float *__restrict__ frame, int start_block, int levels,
                                                               unsigned long long *__restrict__ counts = nullptr) {
  extern __shared__ uint32_t lds[];
  __shared__ float acc[kBlockPixels * 3];
  __shared__ int next_job;
  __shared__ unsigned long long best[kRB];
  __shared__ int ins[kStats ? kRB : 1];
  const BlockLds L = {lds + threadIdx.x, (float *)(lds + levels * kStackWordsPerLevel * kRB), acc, &next_job,
                      best + (threadIdx.x & ~63u), kStats ? ins + (threadIdx.x & ~63u) : nullptr};
  RenderCounters rc;
  render_block<PathT, kStats>(T, tb, P, K, frame, start_block + (int)blockIdx.x, L, PathT::kRefillMin, &rc);
  if (kStats) flush_counters(rc, counts);
}

// ------------------------------------------------------------------ streamed blocks
// k_render_stream: the same path loop with no barrier between framebuffer blocks.  A
// workgroup is persistent and takes blocks from a counter in global memory; its lanes take
// path jobs from one stream that runs through those blocks in turn (job q: pixel q / spp of
// the stream, sample q % spp; stream pixel u is pixel u % 128 of the workgroup's block
// number u / 128), so lanes whose paths end near the end of a block start the next block's
// paths instead of waiting for the block's longest path.  In the one-block-per-workgroup
// kernel those waits took 21% of the path loop's cycles and the barrier after it 7% at
// 1080p, 16 spp.
//   * Sample colours are staged in a ring of `nslots` pixel slots of spp samples
//     (nslots * spp <= kStreamSamples).  Stream pixel u uses slot u % nslots and may start
//     once the slot's owner is u, i.e. once pixel u - nslots has been written out.
//   * The lane that finishes a pixel's last sample sums its spp staged colours in sample
//     order, as render_block does, writes the pixel, and hands the slot to u + nslots.
//     Frames are therefore bit-identical to render_block's.
//   * Block number k of the workgroup is fetched by the lane that takes its first job, and
//     only once block k - 1 has been fetched, so block ids rise with k and the blocks past
//     the launch's range form a suffix of the stream.  The ids sit in a ring of
//     kStreamBlockRing entries; a job of block number k starts only once every pixel of block
//     number k - kStreamBlockRing has been written (pixels written per ring entry are
//     counted cumulatively), so no job still needs an entry when it is reused.
//   * Progress: the oldest pixel not yet written can always start (its slot's previous
//     owner is written, its block's ring entry and the entry before it are fetched), and
//     jobs are taken in stream order, so its jobs are held by lanes; lanes whose job must
//     wait keep retrying while the other lanes work.
constexpr int kStreamSamples = 2048;      // staged sample colours per workgroup
constexpr int kStreamMaxSlots = 1024;     // pixel slots at low sample counts
constexpr int kStreamBlockRing = 16;
constexpr int kStreamMaxSpp = 256;        // higher sample counts take render_block's rounds

__host__ __device__ constexpr int stream_slots(int spp) {
  return (kStreamSamples / spp) < kStreamMaxSlots ? (kStreamSamples / spp) : kStreamMaxSlots;
}

template <class PathT, bool kStats>
__global__ __launch_bounds__(kRB, kRenderWaves) void k_render_stream(DevTree T, Tables tb, dr_render_params P,
                                                                      FrameK K, float *__restrict__ frame, int start_block,
                                                                      int end_block, int levels,
                                                                      unsigned *__restrict__ work,
                                                                      unsigned long long *__restrict__ counts) {
  extern __shared__ uint32_t lds[];
  __shared__ int blocks[kStreamBlockRing];    // (block id + 1) << 8 | stream block number & 255
  __shared__ int px_done[kStreamBlockRing];   // pixels written, cumulative per ring entry
  __shared__ unsigned next_job;
  __shared__ int end_k;                       // first stream block number past the range
  __shared__ unsigned long long best_all[kRB];
  __shared__ int ins_all[kStats ? kRB : 1];
  const int spp = P.fsaa_x * P.fsaa_y;
  const int nslots = stream_slots(spp);
  uint32_t *stk = lds + threadIdx.x;
  float *samp = (float *)(lds + levels * kStackWordsPerLevel * kRB);   // nslots * spp * 3
  int *owner = (int *)(samp + (size_t)nslots * spp * 3);             // nslots
  int *remaining = owner + nslots;                                    // nslots
  unsigned long long *best = best_all + (threadIdx.x & ~63u);
  int *ins = kStats ? ins_all + (threadIdx.x & ~63u) : nullptr;
  for (int t = threadIdx.x; t < nslots; t += kRB) { owner[t] = t; remaining[t] = spp; }
  for (int t = threadIdx.x; t < kStreamBlockRing; t += kRB) { blocks[t] = 0; px_done[t] = 0; }
  if (threadIdx.x == 0) { next_job = 0; end_k = 0x7fffffff; }
  __syncthreads();

  RenderCounters rc;
  CastCounters *cc = kStats ? &rc.cast : nullptr;
  const int nrange = end_block - start_block;
  const uint32_t samp0 = (uint32_t)(P.iteration > 1 ? P.iteration - 1 : 0) * (uint32_t)spp;
  const float spw = uni(1.0f / (((float)P.fsaa_x) * ((float)P.fsaa_y)));
  if constexpr (PathT::kStreamFrameInKernel) K = frame_k_uniform(P);
  const UDiv Dspp = udiv_uniform((uint32_t)spp), Dslots = udiv_uniform((uint32_t)nslots);
  const bool clip = P.cull_backfaces != 0;
  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  PathT ps;
  Walk w;
  w.state = kWalkDone;
  int q = -1;             // stream job of this lane, or -1
  bool live = false;      // the job has started
  bool exhausted = false; // the stream has ended for this lane
  bool nopath = false;    // the job's pixel is outside the image
  int blk = 0;            // the job's block (PathT::kStreamBlockInRegister)
  // stages this lane's sample colour (the lens filter applied: getPixelColor, see
  // render_block); the lane that completes the pixel writes it out
  auto finish = [&](float l0, float l1, float l2) {
    const int u = (int)udiv((uint32_t)q, Dspp), s = q - u * spp, slot = u - (int)udiv((uint32_t)u, Dslots) * nslots;
    float *dst = samp + 3 * ((size_t)slot * spp + s);
    dst[0] = l0;
    dst[1] = l1;
    dst[2] = l2;
    __threadfence_block();
    if (atomicSub(&remaining[slot], 1) == 1) {
      float a[3] = {0.0f, 0.0f, 0.0f};
      const float *src = samp + 3 * (size_t)slot * spp;
      for (int k = 0; k < spp; k++)
        for (int c = 0; c < 3; c++) a[c] += src[3 * k + c] * spw;
      // the job's block (its ring entry stays until every pixel of the block is written)
      const int b = PathT::kStreamBlockInRegister
                        ? blk
                        : (*(volatile int *)&blocks[(u / kBlockPixels) & (kStreamBlockRing - 1)] >> 8) - 1;
      store_pixel(P, K, frame, b, u % kBlockPixels, a);
      remaining[slot] = spp;
      atomicAdd(&px_done[(u / kBlockPixels) & (kStreamBlockRing - 1)], 1);
      __threadfence_block();
      owner[slot] = u + nslots;
    }
    q = -1;
    live = false;
  };
  for (;;) {
    // ---- idle lanes take the next jobs of the stream
    const bool want = (q < 0) && !exhausted;
    const uint64_t wmask = __ballot(want);
    if (wmask && (__popcll(wmask) >= PathT::kStreamRefillMin || __ballot(live) == 0)) {
      const int leader = __ffsll((unsigned long long)wmask) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(&next_job, (unsigned)__popcll(wmask));
      base = __shfl(base, leader);
      if (want) q = (int)(base + (unsigned)__popcll(wmask & below));
    }
    // ---- jobs taken but not started: their block and their pixel slot
    if (q >= 0 && !live) {
      const int u = (int)udiv((uint32_t)q, Dspp), s = q - u * spp;
      const int k = u / kBlockPixels;
      const int kk = k & (kStreamBlockRing - 1);
      if (k >= *(volatile int *)&end_k) {
        q = -1;
        exhausted = true;
      } else if (*(volatile int *)&px_done[kk] >= kBlockPixels * (k / kStreamBlockRing)) {
        // (block number k - kStreamBlockRing is written: the ring entry is this block's)
        int e = *(volatile int *)&blocks[kk];
        if ((e & 0xff) != (k & 0xff) || e == 0) {
          // the first job of block number k fetches its id once k - 1 has one
          const int prev = k > 0 ? *(volatile int *)&blocks[(k - 1) & (kStreamBlockRing - 1)] : 0;
          const bool prev_ok = k == 0 || (prev != 0 && (prev & 0xff) == ((k - 1) & 0xff));
          if (q == k * kBlockPixels * spp && prev_ok) {
            const int id = (int)atomicAdd(work, 1u);
            if (id >= nrange) {
              atomicMin(&end_k, k);
              q = -1;
              exhausted = true;
              e = 0;
            } else {
              e = ((start_block + id + 1) << 8) | (k & 0xff);
              blocks[kk] = e;
            }
          } else {
            e = 0;
          }
        }
        if (e != 0 && (e & 0xff) == (k & 0xff) &&
            *(volatile int *)&owner[u - (int)udiv((uint32_t)u, Dslots) * nslots] == u) {
          live = true;
          if (PathT::kStreamBlockInRegister) blk = (e >> 8) - 1;
          nopath = !ps.start(P, K, (e >> 8) - 1, u % kBlockPixels, s, samp0);
          if (nopath) {
            // a pixel outside the image: a black sample
            if (PathT::kStreamBlackInUpdate) w.state = kWalkDone;
            else finish(0.0f, 0.0f, 0.0f);
          } else {
            walk_begin(T, w, ps.ray_o(), ps.ray_d(), ps.ray_excl(), stk, kRB, nullptr);
            if (kStats) { rc.paths++; rc.segments++; }
          }
        }
      }
    }
    if (__ballot(live) == 0) {
      if (__ballot(q >= 0 || !exhausted) == 0) break;
      if (__ballot(q >= 0) != 0) __builtin_amdgcn_s_sleep(1);   // jobs waiting for a slot or a block id
      continue;
    }
    // ---- one traversal step of every ray in flight
    if (live && w.state == kWalkTrace) walk_step<PathT::kFastSlab>(T, w, stk, kRB, cc);
    // ---- leaf tests, batched
    const bool pend = live && w.state == kWalkLeaf;
    const uint64_t pm = __ballot(pend);
    if (pm && (__popcll(pm) >= kLeafMin || __ballot(live && w.state == kWalkTrace) == 0))
      walk_leaf_wave(T, w, pend, clip, best, ins, cc);
    // ---- path updates of the finished casts, batched
    const uint64_t dm = __ballot(live && w.state == kWalkDone);
    if (dm && (__popcll(dm) >= PathT::kStreamShadeMin || __ballot(live && w.state != kWalkDone) == 0)) {
      if (live && w.state == kWalkDone) {
        if ((PathT::kStreamBlackInUpdate && nopath) || ps.shade(P, K, tb, w)) {
          const float *col = ps.color();
          float c0 = col[0], c1 = col[1], c2 = col[2];
          if (P.lens == DR_SPHERICAL_SEGMENT_LENS) {
            c0 = c0 * P.lens_p[4];
            c1 = c1 * P.lens_p[5];
            c2 = c2 * P.lens_p[6];
          }
          // (a job outside the image stages black; its pixel is never written)
          const bool black = PathT::kStreamBlackInUpdate && nopath;
          finish(black ? 0.0f : c0, black ? 0.0f : c1, black ? 0.0f : c2);
        } else {
          walk_begin(T, w, ps.ray_o(), ps.ray_d(), ps.ray_excl(), stk, kRB, nullptr);
          if (kStats) { rc.segments++; rc.photons += ps.photon() ? 1 : 0; }
        }
      }
    }
  }
  if (kStats) flush_counters(rc, counts);
  // ---- the last wave out resets the block counter for the next launch
  if (lane == 0) {
    __threadfence();
    const unsigned total = gridDim.x * (kRB / 64);
    if (atomicAdd(work + 1, 1u) == total - 1) {
      atomicExch(work, 0u);
      atomicExch(work + 1, 0u);
    }
  }
}

// ------------------------------------------------------------------ adjoint
// Vertex records of each lane's current path ([vertex][word][lane] in the
// workgroup's slice of a scratch buffer): material id, reflection factor A (3),
// direct sum d (3) with L_k = A (d + s L_{k+1}), the sample weight s, and dA/dalbedo
// (1, or 1 - f5 for a Schlick metal, whose A = F0 + (1 - F0) f5).
// The path is replayed from its Philox key; after it ends S_k = d_k + s_k L_{k+1}
// is formed backwards and dL/dalbedo += w dL/dI beta_k dA_k S_k.  dL/demission is
// added during the replay: w dL/dI beta_k A_k (cos_i/pi)(cos_l A/d^2).
// Rays are step-wise walks as in render_block.  The grid is persistent (one vertex-record
// slice per workgroup): a workgroup takes the next framebuffer block from a counter in
// global memory when it has finished one, so blocks of uneven cost spread over the grid as
// hardware dispatch of one workgroup per block would spread them (a fixed assignment g,
// g + gridDim.x, ... measured 20.9 ms against 15.7 ms for one workgroup per block on the
// 512x512 16-spp adjoint).
// Reductions: the lanes of a wave that add to the same parameter in the same path
// update are summed across the wave first (one LDS atomic per distinct material or
// light and channel); the workgroup's LDS sums of one framebuffer block then go to
// the FP64 outputs with one global atomic per parameter, and restart from zero.
constexpr int kVtxWords = 9;

// Sum of v over the 64 lanes of the wave, returned to every lane: an inclusive scan within
// each row of 16 lanes with DPP row shifts (1, 2, 4, 8; a lane without a source adds 0)
// leaves the row totals in lanes 15, 31, 47 and 63, which are read out and added.
template <int kCtrl>
__device__ __forceinline__ float dpp_or_zero(float v) {   // the DPP source's bits, 0 where it has none
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_or_zero<0x111>(v);   // row_shr:1
  v += dpp_or_zero<0x112>(v);   // row_shr:2
  v += dpp_or_zero<0x114>(v);   // row_shr:4
  v += dpp_or_zero<0x118>(v);   // row_shr:8
  return (lane_value(v, 15) + lane_value(v, 31)) + (lane_value(v, 47) + lane_value(v, 63));
}

// Lanes that must share a parameter before their terms are summed across the wave: a sum
// costs six cross-lane steps per channel, a lane's own LDS atomic one contended add.
// (1, 8, 24 and 65 measured 22.8 / 22.5 / 22.2 / 21.8 ms for the 512x512 16-spp adjoint of
// 64 materials and 4 lights)
constexpr int kWaveSumMin = 16;

// Adds v[0..2] of every lane with `has` to acc[3 * key + c], one key at a time: the lanes
// sharing a key are summed across the wave first (one atomic per channel) when there are at
// least kWaveSumMin of them, and add their own terms otherwise.  Wave-uniform call.
__device__ __forceinline__ void wave_add3(float *acc, bool has, int key, const float v[3]) {
  const int lane = (int)__lane_id();
  uint64_t act = __ballot(has);
  if (__popcll(act) < kWaveSumMin) {          // too few terms for any sum across the wave to pay
    if (has) {
      atomicAdd(&acc[3 * key + 0], v[0]);
      atomicAdd(&acc[3 * key + 1], v[1]);
      atomicAdd(&acc[3 * key + 2], v[2]);
    }
    return;
  }
  while (act) {
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int k = __builtin_amdgcn_readlane(key, leader);
    const bool sel = has && key == k;
    if (__popcll(__ballot(sel)) >= kWaveSumMin) {
      const float s0 = wave_sum(sel ? v[0] : 0.0f), s1 = wave_sum(sel ? v[1] : 0.0f), s2 = wave_sum(sel ? v[2] : 0.0f);
      if (lane == leader) {
        atomicAdd(&acc[3 * k + 0], s0);
        atomicAdd(&acc[3 * k + 1], s1);
        atomicAdd(&acc[3 * k + 2], s2);
      }
    } else if (sel) {
      atomicAdd(&acc[3 * k + 0], v[0]);
      atomicAdd(&acc[3 * k + 1], v[1]);
      atomicAdd(&acc[3 * k + 2], v[2]);
    }
    has = has && !sel;
    act = __ballot(has);
  }
}

__global__ __launch_bounds__(kRB, kRenderWaves) void k_backward(DevTree T, const float4 *__restrict__ shade,
                                                                 dr_render_params P, FrameK FK,
                                                                 const float *__restrict__ dl,
                                                                 double *__restrict__ g_albedo,
                                                                 double *__restrict__ g_emission,
                                                                 float *__restrict__ vtx_all,
                                                                 int *__restrict__ next_block, int nblocks,
                                                                 int levels, int maxv) {
  extern __shared__ uint32_t lds[];
  __shared__ int next_job, cur_block;
  __shared__ unsigned long long best_all[kRB];
  unsigned long long *best = best_all + (threadIdx.x & ~63u);
  uint32_t *stk = lds + threadIdx.x;
  float *gacc = (float *)(lds + levels * kStackWordsPerLevel * kRB);   // nmat*3 albedo, then nlights*3 emission
  float *vtx = vtx_all + (size_t)blockIdx.x * maxv * kVtxWords * kRB;
  const int ngrad = 3 * (P.nmat + P.nlights);
  const int spp = P.fsaa_x * P.fsaa_y;
  const uint32_t samp0 = (uint32_t)(P.iteration > 1 ? P.iteration - 1 : 0) * (uint32_t)spp;
  const float spw = uni(1.0f / (((float)P.fsaa_x) * ((float)P.fsaa_y)));
  const UDiv Dspp = udiv_uniform((uint32_t)spp);
  const bool clip = P.cull_backfaces != 0;
  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int t = threadIdx.x; t < ngrad; t += kRB) gacc[t] = 0.0f;
  const int njobs = kBlockPixels * spp;
  auto V_ = [&](int k, int w) -> float & { return vtx[((size_t)k * kVtxWords + w) * kRB + threadIdx.x]; };

  for (;;) {
    if (threadIdx.x == 0) {
      next_job = 0;
      cur_block = atomicAdd(next_block, 1);
    }
    __syncthreads();
    const int block = cur_block;      // read before the barriers that end this iteration
    if (block >= nblocks) break;
    Path st;
    Walk w;
    w.state = kWalkDone;
    int K = 0;                      // vertices recorded for the current path
    bool fin = false;               // the current path has ended; its albedo terms are pending
    float wdl[3] = {0, 0, 0};       // spw * dL/dI (* lens filter) of the current pixel
    bool exhausted = false;
    // the vertex's reflection factor and dA/dalbedo, then the next bounce (or the path's end)
    auto advance = [&]() {
      const dr_material M = P.materials[st.mid];
      float f5 = 0.0f;
      const bool metal = st.kind == DR_METAL && st.has_sample;
      if (metal) f5 = schlick5(st.nu, st.sdir);
      for (int c = 0; c < 3; c++) V_(K - 1, 1 + c) = metal ? M.albedo[c] + (1.0f - M.albedo[c]) * f5 : M.albedo[c];
      V_(K - 1, 8) = metal ? 1.0f - f5 : 1.0f;
      V_(K - 1, 7) = st.has_sample ? st.sweight : 0.0f;
      if (!st.has_sample) { fin = true; return; }
      for (int c = 0; c < 3; c++) st.beta[c] = st.beta[c] * (V_(K - 1, 1 + c) * st.sweight);
      st.o = st.p;
      st.d = st.sdir;
      st.restrict_tri = st.h;
      st.layer++;
      st.phase = 0;
    };
    for (;;) {
      const bool want = (st.job < 0) && !exhausted;
      const uint64_t wmask = __ballot(want);
      // new path jobs once kRenderRefillMin lanes are idle (or no lane is busy), as in render_block
      // (thresholds 1, 4, 8, 16 measured 21.3 / 21.2 / 21.2 / 21.5 ms for the 512x512 adjoint)
      if (wmask && (__popcll(wmask) >= kRenderRefillMin || __ballot(st.job >= 0) == 0)) {
        const int leader = __ffsll((unsigned long long)wmask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&next_job, __popcll(wmask));
        base = __shfl(base, leader);
        if (want) {
          const int slot = base + __popcll(wmask & below);
          if (slot < njobs) {
            st.job = slot;
            const int lp = (int)udiv((uint32_t)slot, Dspp), s = slot - lp * spp;
            K = 0;
            if (!start_path(P, FK, block, lp, s, samp0, st)) {
              st.job = -1;
            } else {
              const float *g = dl + 4 * ((size_t)kBlockPixels * block + lp);
              // SphericalSegmentLense::getPixelColor scales the sample colour by its filter
              for (int c = 0; c < 3; c++)
                wdl[c] = spw * g[c] * (P.lens == DR_SPHERICAL_SEGMENT_LENS ? P.lens_p[4 + c] : 1.0f);
              walk_begin(T, w, st.o, st.d, st.restrict_tri, stk, kRB, nullptr);
            }
          } else {
            exhausted = true;
          }
        }
      }
      if (__ballot(st.job >= 0) == 0) {
        if (__ballot(!exhausted) == 0) break;
        continue;
      }
      if (st.job >= 0 && w.state == kWalkTrace) walk_step(T, w, stk, kRB, nullptr);
      const bool pend = st.job >= 0 && w.state == kWalkLeaf;
      const uint64_t pm = __ballot(pend);
      if (pm && (__popcll(pm) >= kLeafMin || __ballot(st.job >= 0 && w.state == kWalkTrace) == 0))
        walk_leaf_wave(T, w, pend, clip, best, nullptr, nullptr);
      const uint64_t dm = __ballot(st.job >= 0 && w.state == kWalkDone);
      if (dm && (__popcll(dm) >= kShadeMin || __ballot(st.job >= 0 && w.state != kWalkDone) == 0)) {
        bool lit = false;
        float ev[3] = {0.0f, 0.0f, 0.0f};
        int lkey = 0;
        if (st.job >= 0 && w.state == kWalkDone) {
          bool photons = false;
          if (st.phase == 0) {
            if (!w.hit) fin = true;
            else {
              setup_vertex(P, FK, shade, w.h, w.ray.d, st);
              V_(K, 0) = __int_as_float(st.mid);
              for (int c = 0; c < 4; c++) V_(K, 4 + c) = 0.0f;
              K++;
              photons = st.kind == DR_LAMBERT;
              if (!photons) advance();
            }
          } else {
            const V3 e = w.h.p - st.p;
            lit = w.hit && (w.h.id == st.h || (e.x * e.x + e.y * e.y + e.z * e.z) <= (8.0f * kVecEps));
            if (lit) {
              const dr_material M = P.materials[st.mid];
              const float *E = P.lights[st.light].emission;
              for (int c = 0; c < 3; c++) {
                V_(K - 1, 4 + c) += st.gl_k * E[c];
                ev[c] = wdl[c] * st.beta[c] * M.albedo[c] * st.gl_k;
              }
              lkey = P.nmat + st.light;
            }
            photons = true;
          }
          if (photons && !next_photon(P, st)) advance();
          if (!fin) walk_begin(T, w, st.o, st.d, st.restrict_tri, stk, kRB, nullptr);
        }
        // dL/demission of this update's lit photons
        wave_add3(gacc, lit, lkey, ev);
        // dL/dalbedo of the paths that ended: S_k backwards, then beta_k forwards
        if (__ballot(fin)) {
          if (fin) {
            float Ln[3] = {0, 0, 0};
            for (int k = K - 1; k >= 0; k--)
              for (int c = 0; c < 3; c++) {
                const float S = V_(k, 4 + c) + V_(k, 7) * Ln[c];
                Ln[c] = V_(k, 1 + c) * S;
                V_(k, 4 + c) = S;
              }
          }
          float beta[3] = {1, 1, 1};
          int kmax = fin ? K : 0;
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
          for (int k = 0; k < kmax; k++) {
            const bool has = fin && k < K;
            float av[3] = {0.0f, 0.0f, 0.0f};
            int m = 0;
            if (has) {
              m = __float_as_int(V_(k, 0));
              const float dA = V_(k, 8);
              for (int c = 0; c < 3; c++) {
                av[c] = wdl[c] * beta[c] * dA * V_(k, 4 + c);
                beta[c] = beta[c] * (V_(k, 1 + c) * V_(k, 7));
              }
            }
            wave_add3(gacc, has, m, av);
          }
          if (fin) { st.job = -1; fin = false; }
        }
      }
    }
    __syncthreads();
    // this block's sums into the FP64 gradients
    for (int t = threadIdx.x; t < ngrad; t += kRB) {
      const float v = gacc[t];
      if (v != 0.0f) {
        if (t < 3 * P.nmat) atomicAdd(&g_albedo[t], (double)v);
        else atomicAdd(&g_emission[t - 3 * P.nmat], (double)v);
      }
      gacc[t] = 0.0f;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ batched requests
struct DevRequest {
  DevTree T;
  const float4 *shade;
  const float *faces;            // shading records of the reference plugins
  dr_render_params P;
  FrameK K;                      // frame_k(P), formed on the host
  float *frame;
};

__global__ void k_zero(int *counter, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) counter[threadIdx.x] = 0;
}

// Persistent workgroups drain the queue of (request, first block, block count) items
// of the requests of one plugin set.
template <class PathT>
__global__ __launch_bounds__(kRB, kRenderWaves) void k_batch(const DevRequest *__restrict__ reqs, const int4 *__restrict__ items,
                                                int nitems, int *__restrict__ counter, int levels, int refill_min) {
  extern __shared__ uint32_t lds[];
  __shared__ float acc[kBlockPixels * 3];
  __shared__ int next_job;
  __shared__ int item;
  __shared__ unsigned long long best[kRB];
  const BlockLds L = {lds + threadIdx.x, (float *)(lds + levels * kStackWordsPerLevel * kRB), acc, &next_job,
                      best + (threadIdx.x & ~63u), nullptr};
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);   // the same in every lane: scalar loads below
    __syncthreads();
    if (it >= nitems) break;
    const int4 w = items[it];
    const DevRequest &R = reqs[w.x];
    const Tables tb = {R.shade, R.faces};
    for (int b = w.y; b < w.y + w.z; b++) render_block<PathT>(R.T, tb, R.P, R.K, R.frame, b, L, refill_min);
  }
}

static int check_params(dr_scene s, const dr_render_params *p) {
  if (!s || !p) return fail(DR_ERR_ARG, "render: null scene or parameters");
  if (p->shading && p->shading->scene != s) return fail(DR_ERR_ARG, "render: shading tables of another scene");
  if (!p->shading && !s->has_materials)
    return fail(DR_ERR_ARG, "render: no material ids (dr_scene_set_materials or dr_render_params.shading)");
  const int max_mat = p->shading ? p->shading->max_mat : s->max_mat;
  if (p->width < 1 || p->height < 1 || p->fsaa_x < 1 || p->fsaa_y < 1 || p->max_bounces < 0 || p->max_bounces > 64)
    return fail(DR_ERR_ARG, "render: bad image size, subpixel grid or max_bounces");
  if (p->nlights < 0 || p->nlights > kMaxLights || (p->nlights > 0 && !p->lights))
    return fail(DR_ERR_ARG, "render: 0 <= nlights <= 16 with a light table required");
  if (!p->materials || p->nmat <= max_mat) return fail(DR_ERR_ARG, "render: material table smaller than the ids in use");
  if (!(p->rr_survive > 0.0f && p->rr_survive <= 1.0f)) return fail(DR_ERR_ARG, "render: rr_survive must be in (0, 1]");
  if (s->max_depth > kMaxTreeDepth) return fail(DR_ERR_ARG, "render: octree deeper than the traversal stack");
  if (p->lens != DR_STOCHASTIC_LENS && p->lens != DR_PERSPECTIVE_LENS && p->lens != DR_SPHERICAL_SEGMENT_LENS)
    return fail(DR_ERR_ARG, "render: unknown lens");
  if (p->plugins != DR_PLUGINS_BENCHMARK && p->plugins != DR_PLUGINS_REFERENCE)
    return fail(DR_ERR_ARG, "render: unknown plugin set");
  if (p->nmaps < 0 || p->ntextures < 0 || (p->nmaps > 0 && !p->maps) || (p->ntextures > 0 && (!p->textures || !p->texels)))
    return fail(DR_ERR_ARG, "render: texture tables missing");
  return DR_OK;
}

static int block_count(const dr_render_params *p) { return ((p->width + 7) / 8) * ((p->height + 15) / 16); }

// The shading tables a render reads: its dr_shading handle's, or the scene's own.
static Tables tables_of(dr_scene s, const dr_render_params *p) {
  if (p->shading) return {(const float4 *)p->shading->shade.p, (const float *)p->shading->faces.p};
  return {(const float4 *)s->shade.p, (const float *)s->faces.p};
}

}  // namespace dr

namespace dr {
static int launch_render(dr_scene s, const dr_render_params *p, float *frame, int start_block, int end_block,
                         unsigned long long *counts, void *stream) {
  int rc = check_params(s, p);
  if (rc != DR_OK) return rc;
  if (!frame || start_block < 0 || end_block > block_count(p) || start_block > end_block)
    return fail(DR_ERR_ARG, "dr_render_blocks: bad frame or block range");
  if (start_block == end_block) return DR_OK;
  const int levels = s->max_depth > 0 ? s->max_depth : 1;
  const Tables tb = tables_of(s, p);
  const bool ref = p->plugins == DR_PLUGINS_REFERENCE;
  const int spp = p->fsaa_x * p->fsaa_y;
  // streamed blocks (k_render_stream) when each workgroup gets at least two blocks, unless the
  // sample count or the stream length is too large; one block per workgroup otherwise
  const long long nrange = end_block - start_block;
  if (spp <= kStreamMaxSpp && (nrange + 1) * kBlockPixels * spp < (1ll << 31) && end_block < (1 << 23)) {
    const int nslots = stream_slots(spp);
    const size_t lds_s =
        ((size_t)levels * kStackWordsPerLevel * kRB + (size_t)nslots * spp * 3 + 2 * (size_t)nslots) * 4;
    auto kern = counts ? (ref ? k_render_stream<RefPath, true> : k_render_stream<BenchPath, true>)
                       : (ref ? k_render_stream<RefPath, false> : k_render_stream<BenchPath, false>);
    if (lds_s > 64 * 1024)
      DR_HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
    int per_cu = 0;
    DR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kRB, lds_s));
    const long long resident = (long long)s->num_cu * std::max(per_cu, 1);
    const int groups = (int)std::min<long long>(nrange, resident);
    if (nrange >= 2 * resident) {
      unsigned *work = (unsigned *)s->work.p + 2 * (s->work_slot.fetch_add(1) % kWorkSlots);
      hipLaunchKernelGGL(kern, dim3(groups), dim3(kRB), lds_s, (hipStream_t)stream, s->tree, tb, *p, frame_k(*p),
                         frame, start_block, end_block, levels, work, counts);
      DR_HIP_CHECK(hipGetLastError());
      return DR_OK;
    }
  }
  const size_t lds = ((size_t)levels * kStackWordsPerLevel * kRB + (size_t)kBlockPixels * kRoundSamples * 3) * 4;
  auto kernel = counts ? (ref ? k_render<RefPath, true> : k_render<BenchPath, true>)
                       : (ref ? k_render<RefPath, false> : k_render<BenchPath, false>);
  if (lds > 64 * 1024)
    DR_HIP_CHECK(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kernel, dim3(end_block - start_block), dim3(kRB), lds, (hipStream_t)stream, s->tree, tb, *p,
                     frame_k(*p), frame, start_block, levels, counts);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}
}  // namespace dr

extern "C" int dr_render_blocks(dr_scene s, const dr_render_params *p, float *frame, int start_block, int end_block,
                                void *stream) {
  return dr::launch_render(s, p, frame, start_block, end_block, nullptr, stream);
}

extern "C" int dr_render_stats(dr_scene s, const dr_render_params *p, float *frame, int start_block, int end_block,
                               uint64_t *counts, void *stream) {
  if (!counts) return dr::fail(DR_ERR_ARG, "dr_render_stats: null counters");
  return dr::launch_render(s, p, frame, start_block, end_block, (unsigned long long *)counts, stream);
}

namespace dr {

// Launch shape of the adjoint: LDS bytes (traversal stacks + gradient sums) and the size
// of one workgroup's vertex records.
static int backward_shape(dr_scene s, const dr_render_params *p, int &levels, int &maxv, size_t &lds, size_t &slice) {
  levels = s->max_depth > 0 ? s->max_depth : 1;
  maxv = p->max_bounces + 1;
  lds = ((size_t)levels * kStackWordsPerLevel * kRB + 3 * (size_t)(p->nmat + p->nlights)) * 4;
  slice = (size_t)maxv * kVtxWords * kRB * sizeof(float);
  if (lds > 160 * 1024) return fail(DR_ERR_ARG, "dr_backward: too many materials and lights for LDS");
  if (lds > 64 * 1024)
    DR_HIP_CHECK(hipFuncSetAttribute((const void *)k_backward, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return DR_OK;
}

// Workgroups of the persistent adjoint grid: one per resident slot, at most one per
// framebuffer block, and at most as many as `vtx_bytes` of vertex records hold.
static int backward_groups(dr_scene s, size_t lds, size_t slice, int nblocks, size_t vtx_bytes, int &groups) {
  int per_cu = 0;
  DR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_backward, kRB, lds));
  const long long resident = (long long)s->num_cu * std::max(per_cu, 1);
  groups = (int)std::min<long long>({(long long)nblocks, resident, (long long)(vtx_bytes / slice)});
  return DR_OK;
}

static int check_backward(dr_scene s, const dr_render_params *p, const float *dl_dframe, const double *g_albedo,
                          const double *g_emission) {
  int rc = check_params(s, p);
  if (rc != DR_OK) return rc;
  if (!dl_dframe || !g_albedo || (p->nlights > 0 && !g_emission)) return fail(DR_ERR_ARG, "dr_backward: null buffers");
  if (p->plugins != DR_PLUGINS_BENCHMARK) return fail(DR_ERR_ARG, "dr_backward: only the benchmark plugin set has an adjoint");
  return DR_OK;
}

// Scratch of one adjoint launch: the block counter, then `groups` vertex-record slices.
constexpr size_t kCounterBytes = 256;

static void launch_backward(dr_scene s, const dr_render_params *p, const float *dl, double *ga, double *ge,
                            void *scratch_mem, int groups, int nblocks, int levels, int maxv, size_t lds,
                            hipStream_t stream) {
  int *counter = (int *)scratch_mem;
  float *vtx = (float *)((char *)scratch_mem + kCounterBytes);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, counter, 1);
  hipLaunchKernelGGL(k_backward, dim3(groups), dim3(kRB), lds, stream, s->tree, tables_of(s, p).shade, *p, frame_k(*p),
                     dl, ga, ge, vtx, counter, nblocks, levels, maxv);
}

}  // namespace dr

namespace dr {

// Block counter and vertex records of dr_backward_noalloc: a file-scope device array
// (kNoAllocBytes on every device the library runs on, held until the process ends).  Host
// code never passes `scratch` itself (its host-side name is not a device address): it
// passes the address hipGetSymbolAddress gives for the current device, resolved once per
// device.
constexpr size_t kNoAllocBytes = 40u << 20;
static __device__ __align__(256) float scratch[kNoAllocBytes / sizeof(float)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_init(int device, float **out) {
  if (device < 0 || device >= kMaxDevices) return fail(DR_ERR_ARG, "dr_backward_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(DR_ERR_HIP, std::string("dr_backward_noalloc: scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (float *)ptr;
  }
  if (out) *out = g_scratch_ptr[device];
  return DR_OK;
}

}  // namespace dr

extern "C" int dr_backward_noalloc(dr_scene s, const dr_render_params *p, const float *dl_dframe, double *g_albedo,
                                   double *g_emission, void *stream) {
  int rc = dr::check_backward(s, p, dl_dframe, g_albedo, g_emission);
  if (rc != DR_OK) return rc;
  int levels, maxv, groups, device = 0;
  size_t lds, slice;
  if ((rc = dr::backward_shape(s, p, levels, maxv, lds, slice)) != DR_OK) return rc;
  if (slice > dr::kNoAllocBytes - dr::kCounterBytes)
    return dr::fail(DR_ERR_ARG, "dr_backward_noalloc: one workgroup's vertex records exceed the 40 MiB scratch array");
  DR_HIP_CHECK(hipGetDevice(&device));
  float *mem = nullptr;
  if ((rc = dr::noalloc_scratch_init(device, &mem)) != DR_OK) return rc;
  const int nblocks = dr::block_count(p);
  // as many persistent workgroups as the array holds vertex records for (after the block counter)
  if ((rc = dr::backward_groups(s, lds, slice, nblocks, dr::kNoAllocBytes - dr::kCounterBytes, groups)) != DR_OK)
    return rc;
  dr::launch_backward(s, p, dl_dframe, g_albedo, g_emission, mem, groups, nblocks, levels, maxv, lds,
                      (hipStream_t)stream);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}

extern "C" int dr_backward(dr_scene s, const dr_render_params *p, const float *dl_dframe, double *g_albedo,
                           double *g_emission, void *stream) {
  int rc = dr::check_backward(s, p, dl_dframe, g_albedo, g_emission);
  if (rc != DR_OK) return rc;
  int levels, maxv, groups;
  size_t lds, slice;
  if ((rc = dr::backward_shape(s, p, levels, maxv, lds, slice)) != DR_OK) return rc;
  const int nblocks = dr::block_count(p);
  if ((rc = dr::backward_groups(s, lds, slice, nblocks, SIZE_MAX, groups)) != DR_OK) return rc;
  // block counter and vertex records: the scene's buffer for this stream, grown when a call needs more
  void *mem = nullptr;
  if ((rc = dr::scene_scratch(s, (hipStream_t)stream, dr::kCounterBytes + (size_t)groups * slice, &mem)) != DR_OK)
    return rc;
  dr::launch_backward(s, p, dl_dframe, g_albedo, g_emission, mem, groups, nblocks, levels, maxv, lds,
                      (hipStream_t)stream);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}

// ------------------------------------------------------------------ batch host side
struct dr_batch_s {
  int n = 0, levels = 1, num_cu = 1;
  int nitems[2] = {0, 0};          // work items of the benchmark / reference plugin requests
  dr_batch_tuning tuning{1, 2, 1};
  dr::DevBuf reqs, items[2], counter;
  std::vector<dr::DevRequest> host_reqs;
  std::vector<dr_scene> scenes;      // references held until dr_batch_destroy
  std::vector<dr_shading> shadings;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace dr {

static size_t batch_lds(int levels) {
  return ((size_t)levels * kStackWordsPerLevel * kRB + (size_t)kBlockPixels * kRoundSamples * 3) * 4;
}

// Queue of (request, first block, count) of the requests of plugin set `set` (0 =
// benchmark, 1 = reference), heaviest per-block cost first.
static std::vector<int4> make_items(const std::vector<DevRequest> &R, int tile, int set) {
  std::vector<std::pair<double, int4>> v;
  for (int r = 0; r < (int)R.size(); r++) {
    const dr_render_params &P = R[r].P;
    if ((P.plugins == DR_PLUGINS_REFERENCE) != (set == 1)) continue;
    const int nb = block_count(&P);
    const double cost = (double)P.fsaa_x * P.fsaa_y * (1.0 + P.max_bounces) * (1.0 + P.nlights);
    for (int b = 0; b < nb; b += tile) v.push_back({cost * std::min(tile, nb - b), make_int4(r, b, std::min(tile, nb - b), 0)});
  }
  std::stable_sort(v.begin(), v.end(), [](const std::pair<double, int4> &a, const std::pair<double, int4> &b) {
    return a.first > b.first;
  });
  std::vector<int4> out;
  out.reserve(v.size());
  for (auto &e : v) out.push_back(e.second);
  return out;
}

// Queue resets, then one persistent kernel per plugin set that has requests (the
// reference estimator's kernel after the benchmark one, on the same stream).
static int launch_batch(const DevRequest *reqs_dev, const int4 *const items_dev[2], const int nitems[2], int *counter,
                        int levels, const dr_batch_tuning &t, int num_cu, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, counter, 2);
  const dim3 grid(num_cu * t.groups_per_cu);
  if (nitems[0] > 0)
    hipLaunchKernelGGL(k_batch<BenchPath>, grid, dim3(kRB), batch_lds(levels), stream, reqs_dev, items_dev[0],
                       nitems[0], counter, levels, t.refill_min);
  if (nitems[1] > 0)
    hipLaunchKernelGGL(k_batch<RefPath>, grid, dim3(kRB), batch_lds(levels), stream, reqs_dev, items_dev[1],
                       nitems[1], counter + 1, levels, t.refill_min);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}

// Times one tuning on the proxy (every request at 1x1 spp, iteration 1, scratch frames).
static int time_tuning(const std::vector<DevRequest> &proxy, const dr_batch_tuning &t, int levels, int num_cu,
                       hipStream_t stream, DevBuf &reqs_dev, DevBuf items_dev[2], int *counter, float &ms) {
  int nitems[2];
  const int4 *items_p[2];
  int rc = DR_OK;
  for (int set = 0; set < 2 && rc == DR_OK; set++) {
    const std::vector<int4> items = make_items(proxy, t.tile_blocks, set);
    nitems[set] = (int)items.size();
    dev_free(items_dev[set]);
    rc = dev_upload(items_dev[set], items.data(), items.size() * sizeof(int4));
    items_p[set] = (const int4 *)items_dev[set].p;
  }
  if (rc != DR_OK) return rc;
  hipEvent_t e0, e1;
  DR_HIP_CHECK(hipEventCreate(&e0));
  DR_HIP_CHECK(hipEventCreate(&e1));
  rc = launch_batch((const DevRequest *)reqs_dev.p, items_p, nitems, counter, levels, t, num_cu, stream);
  if (rc == DR_OK) {
    DR_HIP_CHECK(hipEventRecord(e0, stream));
    for (int k = 0; k < 2 && rc == DR_OK; k++)
      rc = launch_batch((const DevRequest *)reqs_dev.p, items_p, nitems, counter, levels, t, num_cu, stream);
    DR_HIP_CHECK(hipEventRecord(e1, stream));
    DR_HIP_CHECK(hipEventSynchronize(e1));
    DR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

static int autotune(const std::vector<DevRequest> &R, int levels, int num_cu, hipStream_t stream,
                    dr_batch_tuning &best) {
  std::vector<DevRequest> proxy = R;
  std::vector<DevBuf> frames(R.size());
  int rc = DR_OK;
  for (size_t r = 0; r < R.size() && rc == DR_OK; r++) {
    proxy[r].P.fsaa_x = proxy[r].P.fsaa_y = 1;
    proxy[r].P.iteration = 1;
    proxy[r].K = frame_k(proxy[r].P);
    frames[r].bytes = (size_t)block_count(&R[r].P) * kBlockPixels * 16;
    hipError_t e = hipMalloc(&frames[r].p, frames[r].bytes);
    if (e != hipSuccess) { frames[r].p = nullptr; rc = fail(DR_ERR_ALLOC, "dr_batch_create: scratch frames"); }
    proxy[r].frame = (float *)frames[r].p;
  }
  DevBuf reqs_dev, items_dev[2], counter;
  if (rc == DR_OK) rc = dev_upload(reqs_dev, proxy.data(), proxy.size() * sizeof(DevRequest));
  const int zero[2] = {0, 0};
  if (rc == DR_OK) rc = dev_upload(counter, zero, sizeof(zero));
  best = dr_batch_tuning{1, 2, kRenderRefillMin};
  float best_ms = 0.0f;
  if (rc == DR_OK) rc = time_tuning(proxy, best, levels, num_cu, stream, reqs_dev, items_dev, (int *)counter.p, best_ms);
  // coordinate descent: tile, then persistent groups per CU, then refill cadence (on the
  // 64-request benchmark mix, refills of 1, 4, 8 and 16 idle lanes measured 286.2, 284.8,
  // 283.3 and 283.3 ms; the 1-spp proxy cannot rank values that close, so the candidates
  // stay near the best)
  const int tiles[] = {1, 2, 4, 8}, groups[] = {1, 2, 3, 4}, refills[] = {4, 8, 16};
  for (int axis = 0; axis < 3 && rc == DR_OK; axis++) {
    const int *vals = axis == 0 ? tiles : (axis == 1 ? groups : refills);
    const int nv = axis == 2 ? 3 : 4;
    dr_batch_tuning base = best;
    for (int i = 0; i < nv && rc == DR_OK; i++) {
      dr_batch_tuning t = base;
      int &field = axis == 0 ? t.tile_blocks : (axis == 1 ? t.groups_per_cu : t.refill_min);
      if (field == vals[i]) continue;
      field = vals[i];
      float ms = 0.0f;
      rc = time_tuning(proxy, t, levels, num_cu, stream, reqs_dev, items_dev, (int *)counter.p, ms);
      if (rc == DR_OK && ms < best_ms) { best_ms = ms; best = t; }
    }
  }
  dev_free(reqs_dev);
  dev_free(items_dev[0]);
  dev_free(items_dev[1]);
  dev_free(counter);
  for (auto &f : frames) dev_free(f);
  return rc;
}

}  // namespace dr

extern "C" int dr_batch_create(const dr_request *reqs, int n, const dr_batch_tuning *tuning, dr_batch *out) {
  if (!reqs || n < 1 || !out) return dr::fail(DR_ERR_ARG, "dr_batch_create: bad arguments");
  *out = nullptr;
  dr_batch b = new (std::nothrow) dr_batch_s();
  if (!b) return dr::fail(DR_ERR_ALLOC, "dr_batch_create: out of host memory");
  int rc = DR_OK;
  for (int r = 0; r < n && rc == DR_OK; r++) {
    rc = dr::check_params(reqs[r].scene, &reqs[r].params);
    if (rc == DR_OK && !reqs[r].frame) rc = dr::fail(DR_ERR_ARG, "dr_batch_create: null frame");
    if (rc != DR_OK) break;
    dr::DevRequest d;
    const dr::Tables tb = dr::tables_of(reqs[r].scene, &reqs[r].params);
    d.T = reqs[r].scene->tree;
    d.shade = tb.shade;
    d.faces = tb.faces;
    d.P = reqs[r].params;
    d.K = dr::frame_k(d.P);
    dr::scene_retain(reqs[r].scene);
    b->scenes.push_back(reqs[r].scene);
    if (reqs[r].params.shading) {
      dr::shading_retain(reqs[r].params.shading);
      b->shadings.push_back(reqs[r].params.shading);
    }
    d.frame = reqs[r].frame;
    b->host_reqs.push_back(d);
    b->levels = std::max(b->levels, reqs[r].scene->max_depth);
  }
  if (rc == DR_OK) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) b->num_cu = 1;
    b->n = n;
    if (hipStreamCreate(&b->cap_stream) != hipSuccess) rc = dr::fail(DR_ERR_HIP, "dr_batch_create: stream");
  }
  if (rc == DR_OK) {
    if (tuning) b->tuning = *tuning;
    else rc = dr::autotune(b->host_reqs, b->levels, b->num_cu, b->cap_stream, b->tuning);
  }
  if (rc == DR_OK && (b->tuning.tile_blocks < 1 || b->tuning.groups_per_cu < 1 || b->tuning.refill_min < 1))
    rc = dr::fail(DR_ERR_ARG, "dr_batch_create: tuning values must be >= 1");
  if (rc == DR_OK) {
    rc = dr::dev_upload(b->reqs, b->host_reqs.data(), b->host_reqs.size() * sizeof(dr::DevRequest));
    for (int set = 0; set < 2 && rc == DR_OK; set++) {
      const std::vector<int4> items = dr::make_items(b->host_reqs, b->tuning.tile_blocks, set);
      b->nitems[set] = (int)items.size();
      rc = dr::dev_upload(b->items[set], items.data(), items.size() * sizeof(int4));
    }
    const int zero[2] = {0, 0};
    if (rc == DR_OK) rc = dr::dev_upload(b->counter, zero, sizeof(zero));
  }
  if (rc == DR_OK) {
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const int4 *items_p[2] = {(const int4 *)b->items[0].p, (const int4 *)b->items[1].p};
      rc = dr::launch_batch((const dr::DevRequest *)b->reqs.p, items_p, b->nitems, (int *)b->counter.p, b->levels,
                            b->tuning, b->num_cu, b->cap_stream);
      hipError_t e2 = hipStreamEndCapture(b->cap_stream, &b->graph);
      if (rc == DR_OK && e2 != hipSuccess) rc = dr::fail(DR_ERR_HIP, "dr_batch_create: graph capture");
      if (rc == DR_OK && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = dr::fail(DR_ERR_HIP, "dr_batch_create: graph instantiation");
    } else {
      rc = dr::fail(DR_ERR_HIP, std::string("dr_batch_create: capture: ") + hipGetErrorString(e));
    }
  }
  if (rc != DR_OK) { dr_batch_destroy(b); return rc; }
  *out = b;
  return DR_OK;
}

extern "C" int dr_batch_run(dr_batch b, void *stream) {
  if (!b || !b->exec) return dr::fail(DR_ERR_ARG, "dr_batch_run: bad batch");
  DR_HIP_CHECK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return DR_OK;
}

extern "C" int dr_batch_get_tuning(dr_batch b, dr_batch_tuning *out) {
  if (!b || !out) return dr::fail(DR_ERR_ARG, "dr_batch_get_tuning: bad arguments");
  *out = b->tuning;
  return DR_OK;
}

extern "C" void dr_batch_destroy(dr_batch b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  dr::dev_free(b->reqs);
  dr::dev_free(b->items[0]);
  dr::dev_free(b->items[1]);
  dr::dev_free(b->counter);
  for (dr_shading h : b->shadings) dr::shading_release(h);
  for (dr_scene sc : b->scenes) dr::scene_release(sc);
  delete b;
}
