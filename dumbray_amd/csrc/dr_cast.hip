// Batch closest-hit ray cast.
//
// Each 256-lane block owns a contiguous chunk of rays.  Lanes whose ray has
// finished take the next ray of the chunk: the wave ballots its idle lanes,
// one lane reserves that many rays with a single LDS atomic, and each idle lane
// takes slot base + (number of idle lanes below it) (wave64 ballot + mbcnt
// prefix).  Every loop iteration a lane either advances to its next non-empty
// leaf or tests the triangles of the leaf it holds, so the triangle tests of a
// wave run together instead of interleaving with node expansion.
#include "dr_common.h"

namespace dr {

constexpr int kCastBlock = 256;
constexpr int kCastChunk = 2048;      // rays per block

__device__ __forceinline__ void write_result(int r, bool hit, const Hit &h, int *__restrict__ out_id,
                                             float *__restrict__ out_t, float *__restrict__ out_p) {
  out_id[r] = hit ? h.id : -1;
  out_t[r] = hit ? h.t : FLT_MAX;
  if (out_p) {
    out_p[3 * (size_t)r + 0] = hit ? h.p.x : 0.0f;
    out_p[3 * (size_t)r + 1] = hit ? h.p.y : 0.0f;
    out_p[3 * (size_t)r + 2] = hit ? h.p.z : 0.0f;
  }
}

// Dynamic LDS: `levels` stack words per lane, [level][lane].  kStats: also write
// the per-ray work counters (4 x int32 per ray) to `stats`.
template <bool kStats>
__global__ __launch_bounds__(kCastBlock) void k_cast(DevTree T, const float *__restrict__ rays, int n, int clip,
                                                      const int *__restrict__ exclude, int *__restrict__ out_id,
                                                      float *__restrict__ out_t, float *__restrict__ out_p,
                                                      int *__restrict__ stats) {
  extern __shared__ uint32_t stk_base[];
  __shared__ int next_ray;
  uint32_t *stk = stk_base + threadIdx.x;
  const int chunk_end = min(n, (int)(blockIdx.x + 1) * kCastChunk);
  if (threadIdx.x == 0) next_ray = blockIdx.x * kCastChunk;
  __syncthreads();

  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int r = -1;
  bool exhausted = false;
  TRay ray;
  int ex = -1, sp = -1;
  bool leaf = false;
  int lg = -1, lk = 0, llink = 0, lcnt = 0;
  Hit h;
  h.id = -1; h.t = FLT_MAX; h.p = v3(0.f, 0.f, 0.f);
  CastCounters cnt_local = {0, 0, 0, 0};
  CastCounters *cc = kStats ? &cnt_local : nullptr;
  auto flush_stats = [&](int ray_index) {
    if (kStats) {
      stats[4 * (size_t)ray_index + 0] = cnt_local.expand;
      stats[4 * (size_t)ray_index + 1] = cnt_local.leaves;
      stats[4 * (size_t)ray_index + 2] = cnt_local.plane;
      stats[4 * (size_t)ray_index + 3] = cnt_local.inside;
      cnt_local = {0, 0, 0, 0};
    }
  };

  for (;;) {
    // ---- refill idle lanes from the block's chunk
    const bool want = (r < 0) && !exhausted;
    const uint64_t wmask = __ballot(want);
    if (wmask) {
      const int leader = __ffsll((unsigned long long)wmask) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&next_ray, __popcll(wmask));
      base = __shfl(base, leader);
      if (want) {
        const int slot = base + __popcll(wmask & below);
        if (slot < chunk_end) {
          r = slot;
          const float2 *R = (const float2 *)(rays + 6 * (size_t)r);
          const float2 q0 = R[0], q1 = R[1], q2 = R[2];
          ray = make_tray(v3(q0.x, q0.y, q1.x), v3(q1.y, q2.x, q2.y));
          ex = exclude ? exclude[r] : -1;
          const int st = tree_begin(T, ray, stk, sp, cc);
          if (st == 0) { write_result(r, false, h, out_id, out_t, out_p); flush_stats(r); r = -1; }
          else if (st == 2) { leaf = true; lg = -1; llink = T.root_start; lcnt = T.root_count; }
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(r >= 0) == 0) {
      if (__ballot(!exhausted) == 0) break;
      continue;
    }
    // ---- advance to the next non-empty leaf in DFS order
    if (r >= 0 && !leaf) {
      if (next_leaf(T.groups, T.margin, ray, stk, kCastBlock, sp, lg, lk, llink, lcnt, cc)) leaf = true;
      else { write_result(r, false, h, out_id, out_t, out_p); flush_stats(r); r = -1; }
    }
    // ---- castInLeaf
    if (leaf) {
      leaf = false;
      Box lb;
      if (lg < 0) { lb.s = v3(T.rsx, T.rsy, T.rsz); lb.e = v3(T.rex, T.rey, T.rez); }
      else lb = child_box(T.groups, lg, lk);
      const bool hit = leaf_cast(T.recs, llink, lcnt, lb, ray.o, ray.d, clip != 0, ex, h, cc);
      if (hit || lg < 0) { write_result(r, hit, h, out_id, out_t, out_p); flush_stats(r); r = -1; }
    }
  }
}

static int launch_cast(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
                       float *out_t, float *out_point, int32_t *stats, void *stream) {
  if (s->max_depth > kMaxTreeDepth) return fail(DR_ERR_ARG, "dr_cast: octree deeper than the traversal stack");
  if (n == 0) return DR_OK;
  const int levels = s->max_depth > 0 ? s->max_depth : 1;
  const size_t lds = (size_t)levels * kCastBlock * sizeof(uint32_t);
  const int blocks = (n + kCastChunk - 1) / kCastChunk;
  const int clip = (int)(flags & DR_CAST_CLIP_BACKFACES);
  if (stats)
    hipLaunchKernelGGL(k_cast<true>, dim3(blocks), dim3(kCastBlock), lds, (hipStream_t)stream, s->tree, rays, n, clip,
                       (const int *)exclude, (int *)out_id, out_t, out_point, (int *)stats);
  else
    hipLaunchKernelGGL(k_cast<false>, dim3(blocks), dim3(kCastBlock), lds, (hipStream_t)stream, s->tree, rays, n,
                       clip, (const int *)exclude, (int *)out_id, out_t, out_point, (int *)nullptr);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}

}  // namespace dr

extern "C" int dr_cast(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
                       float *out_t, float *out_point, void *stream) {
  if (!s || n < 0 || (n > 0 && (!rays || !out_id || !out_t))) return dr::fail(DR_ERR_ARG, "dr_cast: bad arguments");
  return dr::launch_cast(s, rays, n, flags, exclude, out_id, out_t, out_point, nullptr, stream);
}

extern "C" int dr_cast_stats(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude,
                             int32_t *out_id, float *out_t, int32_t *out_counts, void *stream) {
  if (!s || n < 0 || (n > 0 && (!rays || !out_id || !out_t || !out_counts)))
    return dr::fail(DR_ERR_ARG, "dr_cast_stats: bad arguments");
  return dr::launch_cast(s, rays, n, flags, exclude, out_id, out_t, nullptr, out_counts, stream);
}
