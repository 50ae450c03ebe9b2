// wave.hpp -- wave64 building blocks for gfx950 kernels (DPP scans, ballots,
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

}  // namespace hs
