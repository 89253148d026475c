// Batch closest-hit ray cast: one ray per lane, wave64, LDS traversal stack.
#include "dr_common.h"

namespace dr {

constexpr int kCastBlock = 256;

// Traversal stack in LDS, [level][lane]: kMaxTreeDepth levels x 256 lanes x 4 B = 24 KiB.
__global__ __launch_bounds__(kCastBlock) void k_cast(DevTree T, const float *__restrict__ rays, int n, int clip,
                                                      const int *__restrict__ exclude, int *__restrict__ out_id,
                                                      float *__restrict__ out_t, float *__restrict__ out_p) {
  __shared__ uint32_t stk[kMaxTreeDepth * kCastBlock];
  const int r = blockIdx.x * kCastBlock + threadIdx.x;
  if (r >= n) return;
  const float2 *R = (const float2 *)(rays + 6 * (size_t)r);
  const float2 q0 = R[0], q1 = R[1], q2 = R[2];
  const V3 o = v3(q0.x, q0.y, q1.x), d = v3(q1.y, q2.x, q2.y);
  const int ex = exclude ? exclude[r] : -1;
  Hit h;
  const bool hit = tree_cast(T, o, d, clip != 0, ex, stk + threadIdx.x, kCastBlock, h);
  out_id[r] = hit ? h.id : -1;
  out_t[r] = hit ? h.t : FLT_MAX;
  if (out_p) {
    out_p[3 * (size_t)r + 0] = hit ? h.p.x : 0.0f;
    out_p[3 * (size_t)r + 1] = hit ? h.p.y : 0.0f;
    out_p[3 * (size_t)r + 2] = hit ? h.p.z : 0.0f;
  }
}

}  // namespace dr

extern "C" int dr_cast(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
                       float *out_t, float *out_point, void *stream) {
  if (!s || n < 0 || (n > 0 && (!rays || !out_id || !out_t))) return dr::fail(DR_ERR_ARG, "dr_cast: bad arguments");
  if (s->max_depth > dr::kMaxTreeDepth) return dr::fail(DR_ERR_ARG, "dr_cast: octree deeper than the traversal stack");
  if (n == 0) return DR_OK;
  const int blocks = (n + dr::kCastBlock - 1) / dr::kCastBlock;
  hipLaunchKernelGGL(dr::k_cast, dim3(blocks), dim3(dr::kCastBlock), 0, (hipStream_t)stream, s->tree, rays, n,
                     (int)(flags & DR_CAST_CLIP_BACKFACES), (const int *)exclude, (int *)out_id, out_t, out_point);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}
