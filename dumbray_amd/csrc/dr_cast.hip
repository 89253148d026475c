// Batch closest-hit ray cast (Octree::cast over a ray batch, Octree.impl.cuh:174-212).
//
// A persistent grid of 256-lane blocks; each lane carries one ray and its own
// DFS stack in LDS.  The loop advances every traversing lane by one bounded step
// per iteration and batches the rarer, divergent work (leaf tests, taking new
// rays) until enough lanes need it; a leaf batch spreads the records of all
// pending leaves over the whole wave (walk_leaf_wave).
#include <algorithm>
#include "dr_common.h"

namespace dr {

constexpr int kCastBlock = 256;
constexpr int kWavesPerBlock = kCastBlock / 64;
constexpr int kWaveChunk = 128;       // rays a wave takes from the global counter at a time
// idle lanes that trigger a refill (incoherent batch: 12, 16, 20, 24, 28, 32 measured 2.16,
// 2.11-2.12, 2.09, 2.07-2.08, 2.09, 2.11 ms)
constexpr int kRefillMin = 24;
constexpr int kLeafMin = 16;          // pending leaves that trigger the (wave-shared) leaf tests

// One bounded step per loop iteration.  A lane that is traversing pops one child:
// an interior child is expanded and pushed, a leaf child is held as pending.
// Pending leaves are tested together once at least kLeafMin lanes hold one (or
// no lane is still traversing), and idle lanes take new rays once kRefillMin of
// them are idle (or no lane is busy), so the node expansions, which dominate the
// work, run with most lanes of the wave active.
//
// Rays are handed out dynamically: a wave owns a range of kWaveChunk rays taken
// from the global counter work[0]; lanes that need a ray take the next ones of the
// range, and the wave takes a new range when it runs out.  The last wave to exit
// (counted in work[1]) sets both counters back to zero for the next launch.
//
// Results are staged per range in LDS (kSlots ranges per wave) and written out as
// whole lines once no ray of the range is still being traced: rays end at scattered
// times, and 4-B stores issued as they end would leave L2 as partial sectors.  When a
// wave needs a slot for a new range while both hold unfinished ranges, the older one
// is written out as far as it is done, and its remaining rays store their results
// directly.
constexpr int kCastCounters = 6;     // per-ray counters of dr_cast_stats
constexpr int kSlots = 2;
constexpr int kPendingId = -2;        // staged entry of a ray still being traced

template <bool kStats>
__global__ __launch_bounds__(kCastBlock, 4) void k_cast(DevTree T, const float *__restrict__ rays, int n, int clip,
                                                      const int *__restrict__ exclude, int *__restrict__ out_id,
                                                      float *__restrict__ out_t, float *__restrict__ out_p,
                                                      int *__restrict__ stats, unsigned *__restrict__ work) {
  extern __shared__ uint32_t stk_base[];
  __shared__ unsigned long long best_all[kCastBlock];   // per lane: best (t, record) key of its leaf
  __shared__ int ins_all[kStats ? kCastBlock : 1];      // per lane: inside tests of its leaf (measurement)
  __shared__ int res_id_all[kWavesPerBlock][kSlots][kWaveChunk];
  __shared__ float res_t_all[kWavesPerBlock][kSlots][kWaveChunk];
  uint32_t *stk = stk_base + threadIdx.x;
  unsigned long long *best = best_all + (threadIdx.x & ~63u);
  int *ins = kStats ? ins_all + (threadIdx.x & ~63u) : nullptr;
  int *res_id = &res_id_all[threadIdx.x >> 6][0][0];
  float *res_t = &res_t_all[threadIdx.x >> 6][0][0];
  constexpr int S = kCastBlock;

  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned wcur = 0, wend = 0;          // the wave's current range of rays (wave-uniform)
  // staging slots (wave-uniform): slot k holds the results of range [base_k, base_k + len_k)
  // while len_k > 0; `cur` is the slot of the current range
  unsigned base0 = 0, base1 = 0;
  int len0 = 0, len1 = 0, cur = 0;
  int r = -1;
  bool exhausted = false;
  Walk w;
  w.state = kWalkDone;
  CastCounters cnt_local = {0, 0, 0, 0, 0, 0};
  CastCounters *cc = kStats ? &cnt_local : nullptr;
  auto finish = [&]() {
    const int id = w.hit ? w.h.id : -1;
    const float t = w.hit ? w.h.t : FLT_MAX;
    const unsigned u = (unsigned)r;
    if (len0 > 0 && u - base0 < (unsigned)len0) {
      res_id[u - base0] = id;
      res_t[u - base0] = t;
    } else if (len1 > 0 && u - base1 < (unsigned)len1) {
      res_id[kWaveChunk + (u - base1)] = id;
      res_t[kWaveChunk + (u - base1)] = t;
    } else {
      out_id[r] = id;
      out_t[r] = t;
    }
    if (out_p) {
      out_p[3 * (size_t)r + 0] = w.hit ? w.h.p.x : 0.0f;
      out_p[3 * (size_t)r + 1] = w.hit ? w.h.p.y : 0.0f;
      out_p[3 * (size_t)r + 2] = w.hit ? w.h.p.z : 0.0f;
    }
    if (kStats) {
      int *o = stats + kCastCounters * (size_t)r;
      o[0] = cnt_local.expand;
      o[1] = cnt_local.leaves;
      o[2] = cnt_local.plane;
      o[3] = cnt_local.inside;
      o[4] = cnt_local.live;
      o[5] = cnt_local.slab;
      cnt_local = {0, 0, 0, 0, 0, 0};
    }
    r = -1;
  };
  // writes the finished entries of slot k (range base, len) as whole lines (wave-uniform)
  auto flush = [&](int k, unsigned base, int len) {
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < len; j += 64) {
      const int id = res_id[k * kWaveChunk + j];
      if (id != kPendingId) {
        out_id[base + j] = id;
        out_t[base + j] = res_t[k * kWaveChunk + j];
      }
    }
  };
  // a range whose rays have all been handed out and finished leaves its slot
  auto retire = [&](int k, unsigned base, int &len) {
    if (len > 0 && (k != cur || wcur == wend) &&
        __ballot(r >= 0 && (unsigned)r - base < (unsigned)len) == 0) {
      flush(k, base, len);
      len = 0;
    }
  };

  for (;;) {
    retire(0, base0, len0);
    retire(1, base1, len1);
    // ---- refill idle lanes
    const bool want = (r < 0) && !exhausted;
    const uint64_t wmask = __ballot(want);
    if (wmask && (__popcll(wmask) >= kRefillMin || __ballot(r >= 0) == 0)) {
      const unsigned need = (unsigned)__popcll(wmask);
      const unsigned idx = (unsigned)__popcll(wmask & below);
      const unsigned rem = wend - wcur;
      int slot;
      if (rem >= need) {
        slot = (int)(wcur + idx);
        wcur += need;
      } else {
        const int leader = __ffsll((unsigned long long)wmask) - 1;
        unsigned nb = 0;
        if (lane == leader) nb = atomicAdd(work, (unsigned)kWaveChunk);
        nb = (unsigned)__builtin_amdgcn_readlane((int)nb, leader);   // wave-uniform: kept in scalar registers
        const unsigned top = min(nb + (unsigned)kWaveChunk, (unsigned)n);
        const unsigned len = nb < top ? top - nb : 0u;
        slot = idx < rem ? (int)(wcur + idx) : (idx - rem < len ? (int)(nb + idx - rem) : -1);
        wcur = nb + min(need - rem, len);
        wend = nb + len;
        if (len > 0) {
          // the new range takes the slot the current range does not hold; an unfinished
          // range still there is written out as far as it is done
          const int k = (cur == 0 ? len0 : len1) > 0 ? 1 - cur : cur;
          if (k == 0 && len0 > 0) flush(0, base0, len0);
          if (k == 1 && len1 > 0) flush(1, base1, len1);
          __builtin_amdgcn_wave_barrier();
          for (int j = lane; j < kWaveChunk; j += 64) res_id[k * kWaveChunk + j] = kPendingId;
          if (k == 0) { base0 = nb; len0 = (int)len; }
          else { base1 = nb; len1 = (int)len; }
          cur = k;
        }
      }
      if (want) {
        if (slot >= 0) {
          r = slot;
          const float2 *R = (const float2 *)(rays + 6 * (size_t)r);
          const float2 q0 = R[0], q1 = R[1], q2 = R[2];
          walk_begin(T, w, v3(q0.x, q0.y, q1.x), v3(q1.y, q2.x, q2.y), exclude ? exclude[r] : -1, stk, S, cc);
          if (w.state == kWalkDone) finish();
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(r >= 0) == 0) {
      if (__ballot(!exhausted) == 0) break;
      continue;
    }
    // ---- one traversal step
    if (r >= 0 && w.state == kWalkTrace) {
      walk_step(T, w, stk, S, cc);
      if (w.state == kWalkDone) finish();
    }

    // ---- castInLeaf for the pending leaves, the whole wave sharing their records
    const bool pend = r >= 0 && w.state == kWalkLeaf;
    const uint64_t pm = __ballot(pend);
    if (pm && (__popcll(pm) >= kLeafMin || __ballot(r >= 0 && w.state == kWalkTrace) == 0)) {
      walk_leaf_wave<true>(T, w, pend, clip != 0, best, ins, cc);
      if (pend && w.state == kWalkDone) finish();
    }
  }
  // every ray of the wave has ended: the ranges still staged go out
  if (len0 > 0) flush(0, base0, len0);
  if (len1 > 0) flush(1, base1, len1);
  // ---- the last wave out resets the counters
  if (lane == 0) {
    __threadfence();
    const unsigned total = gridDim.x * (kCastBlock / 64);
    if (atomicAdd(work + 1, 1u) == total - 1) {
      atomicExch(work, 0u);
      atomicExch(work + 1, 0u);
    }
  }
}

static int launch_cast(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
                       float *out_t, float *out_point, int32_t *stats, void *stream) {
  if (s->max_depth > kMaxTreeDepth) return fail(DR_ERR_ARG, "dr_cast: octree deeper than the traversal stack");
  if (n == 0) return DR_OK;
  const int levels = s->max_depth > 0 ? s->max_depth : 1;
  const int clip = (int)(flags & DR_CAST_CLIP_BACKFACES);
  const size_t lds = (size_t)levels * kStackWordsPerLevel * kCastBlock * sizeof(uint32_t);
  int per_cu = 0;
  DR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cast<false>, kCastBlock, lds));
  // persistent grid: at most one block per resident slot, fewer for small batches
  const long long need = ((long long)n + kWavesPerBlock * kWaveChunk - 1) / (kWavesPerBlock * kWaveChunk);
  const int blocks = (int)std::max(1LL, std::min<long long>(need, (long long)s->num_cu * std::max(per_cu, 1)));
  unsigned *work = (unsigned *)s->work.p + 2 * (s->work_slot.fetch_add(1) % kWorkSlots);
  if (stats)
    hipLaunchKernelGGL(k_cast<true>, dim3(blocks), dim3(kCastBlock), lds, (hipStream_t)stream, s->tree, rays, n, clip,
                       (const int *)exclude, (int *)out_id, out_t, out_point, (int *)stats, work);
  else
    hipLaunchKernelGGL(k_cast<false>, dim3(blocks), dim3(kCastBlock), lds, (hipStream_t)stream, s->tree, rays, n,
                       clip, (const int *)exclude, (int *)out_id, out_t, out_point, (int *)nullptr, work);
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
