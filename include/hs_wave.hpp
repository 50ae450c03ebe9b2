// hs_wave.hpp -- wave64 building blocks for gfx950 kernels (DPP scans, ballots,
// wave-aggregated queue pushes).  All functions must be called by every lane of
// the wave at the same point (convergent).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.hpp"

namespace hs {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// inclusive wave64 prefix sum: DPP row_shr 1,2,4,8 within rows of 16, then
// row_bcast:15 and row_bcast:31 across rows
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_scan(v), 63);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated append: one atomic per wave (Worklist2::setup_push_warp_one,
// LONESTAR/rt/include/worklist.h:350-369, for 64-lane waves).
__device__ __forceinline__ void wave_push(uint32_t *q, uint32_t *cnt, bool want, uint32_t v) {
  const uint64_t m = ballot(want);
  if (m == 0) return;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane_id() == leader) base = add_relaxed(cnt, (uint32_t)__builtin_popcountll(m));
  base = __shfl(base, (int)leader);
  if (want) q[base + lanes_below(m)] = v;
}

// owner lane of edge slot `e` given each lane's inclusive prefix `incl` of a
// non-decreasing sequence: the number of lanes whose prefix is <= e
__device__ __forceinline__ uint32_t wave_owner(uint32_t incl, uint32_t e) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t step = 32; step >= 1; step >>= 1) {
    const uint32_t probe = __shfl(incl, (int)(pos + step - 1));
    if (probe <= e) pos += step;
  }
  return pos;
}


// Segmented gather-sums of 64 runs, one per lane: lane j brings the start `eb`
// and length `len` of its run in `idx` and returns sum(val[idx[e]]) over
// e in [eb, eb + len).  The runs are walked as one flattened sequence, 64
// positions per step (coalesced index loads, the owner lane of each position
// found by wave_owner); STEPS steps per pass have all their index loads and
// gathers in flight before the first segmented wave scan, and each lane adds
// its step sums in step order, so the result depends on the runs alone.
template <int STEPS>
__device__ __forceinline__ float wave_run_gather_sum(const uint32_t *idx, const float *val, uint32_t eb,
                                                     uint32_t len) {
  const uint32_t lane = lane_id();
  const uint32_t incl = wave_incl_scan(len);
  const uint32_t excl = incl - len;
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  const uint32_t base = eb - excl;  // index position of flattened position q of this lane's run
  float mine = 0.f;
  for (uint32_t s0 = 0; s0 < total; s0 += 64 * STEPS) {
    uint32_t oex[STEPS];
    uint32_t col[STEPS];
    float x[STEPS];
#pragma unroll
    for (int u = 0; u < STEPS; ++u) {
      const uint32_t pos = s0 + 64 * u + lane;
      const bool act = pos < total;
      const uint32_t own = wave_owner(incl, act ? pos : total - 1);
      oex[u] = __shfl(excl, (int)own);
      const uint32_t obase = __shfl(base, (int)own);
      col[u] = act ? idx[obase + pos] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < STEPS; ++u) x[u] = col[u] != 0xFFFFFFFFu ? val[col[u]] : 0.f;
#pragma unroll
    for (int u = 0; u < STEPS; ++u) {
      const uint32_t s = s0 + 64 * u;
      if (s >= total) break;
      const uint32_t pos = s + lane;
      float v = x[u];
      // segmented inclusive scan: lane - d is in the same run iff pos - d >= oex
#pragma unroll
      for (uint32_t d = 1; d < 64; d <<= 1) {
        const float up = __shfl_up(v, d);
        if (lane >= d && pos - d >= oex[u]) v += up;
      }
      // lane j takes the scanned value at the last position of its run in this step
      const bool touches = len && excl < s + 64 && incl > s;
      const uint32_t last = (incl < s + 64 ? incl : s + 64) - 1 - s;
      const float got = __shfl(v, (int)(touches ? last : 0u));
      if (touches) mine += got;
    }
  }
  return mine;
}

}  // namespace hs
