// hs_worklist.hpp -- device-side frontier queues and work expansion for
// persistent graph kernels on gfx950 (header-only, HIP).
//
// The role of Worklist2 / PipeContextT (LONESTAR/rt/include/worklist.h:49-585,
// LONESTAR/rt/include/pipe.h:29-123) and of the NP expansion
// (LONESTAR/rt/include/internal.h:23-158), shaped for this device:
//  - a queue counter is one address, and every device-scope atomic on it
//    serializes for ~11 ns across the whole GPU, so appends are aggregated per
//    wave and per batch of candidates (one atomic per wave per batch);
//  - a level-synchronous kernel rotates three counters (input, output, the one
//    to clear for the level after), so the clear never races the readers;
//  - vertices of high degree are not expanded by one workgroup: they go to a
//    chunk list (entries v | chunk << 32) that every wave of the grid drains.
// All wave functions must be called by every lane of the wave (convergent).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.hpp"
#include "hs_wave.hpp"

namespace hs {

// Appends the lanes' candidates v[x] for which want[x] holds (B per lane) to q,
// with one atomic on *cnt per wave for the whole batch.
template <int B>
__device__ __forceinline__ void wave_push_batch(uint32_t *q, uint32_t *cnt, const bool (&want)[B],
                                                const uint32_t (&v)[B]) {
  uint64_t mask[B];
  uint32_t total = 0;
#pragma unroll
  for (int x = 0; x < B; ++x) {
    mask[x] = ballot(want[x]);
    total += (uint32_t)__builtin_popcountll(mask[x]);
  }
  if (total == 0) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = add_relaxed(cnt, total);
  base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
  for (int x = 0; x < B; ++x) {
    if (mask[x] >> lane_id() & 1) q[base + lanes_below(mask[x])] = v[x];
    base += (uint32_t)__builtin_popcountll(mask[x]);
  }
}

// Appends nc[x] chunk entries v[x] | c << 32 (c < nc[x]) per lane, one atomic.
template <int B>
__device__ __forceinline__ void wave_push_chunks_batch(uint64_t *hq, uint32_t *hcnt, const uint32_t (&nc)[B],
                                                       const uint32_t (&v)[B]) {
  uint32_t ncl = 0;
#pragma unroll
  for (int x = 0; x < B; ++x) ncl += nc[x];
  const uint32_t incl = wave_incl_scan(ncl);
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  if (total == 0) return;
  uint32_t base = 0;
  if (lane_id() == 63) base = add_relaxed(hcnt, total);
  base = __builtin_amdgcn_readlane(base, 63) + incl - ncl;
#pragma unroll
  for (int x = 0; x < B; ++x)
    for (uint32_t c = 0; c < nc[x]; ++c) hq[base++] = v[x] | ((uint64_t)c << 32);
}

// Double-buffered frontier of a level-synchronous kernel.  Level k reads
// items[k & 1] (count counts[k % 3]) and appends to items[(k + 1) & 1] (count
// counts[(k + 1) % 3]); its grab counter is grabs[k % 3].  Counters sit
// `stride` words apart (their own cache lines).
struct Frontier {
  uint32_t *items;   // [2][capacity]
  uint32_t *counts;  // 3 counters
  uint32_t *grabs;   // 3 grab counters
  uint32_t capacity;
  uint32_t stride;

  __device__ uint32_t *count_at(uint32_t i) const { return counts + (i % 3) * stride; }
  __device__ uint32_t *grab_at(uint32_t i) const { return grabs + (i % 3) * stride; }
  __device__ const uint32_t *in(uint32_t k) const { return items + (size_t)(k & 1) * capacity; }
  __device__ uint32_t *out(uint32_t k) const { return items + (size_t)((k + 1) & 1) * capacity; }
  __device__ uint32_t in_count(uint32_t k) const { return ld_relaxed(count_at(k)); }
  __device__ uint32_t *out_count(uint32_t k) const { return count_at(k + 1); }
  // one thread of the grid, during level k: clear what level k + 1 will use
  // (its output count, last read at level k - 1, and its grab counter)
  __device__ void clear_for_next(uint32_t k) const {
    st_relaxed(count_at(k + 2), 0u);
    st_relaxed(grab_at(k + 1), 0u);
  }
  // master thread: the first of the next n input items this workgroup takes
  // (work stealing over the level's input)
  __device__ uint32_t grab(uint32_t k, uint32_t n) const { return add_relaxed(grab_at(k), n); }
};

// Fine-grained expansion of up to 64 vertices, one per lane (u, with its
// adjacency [eb, eb + deg)): their edges are walked as one flattened sequence,
// 64 per step, the owner lane of each position found by wave_owner; every B
// steps the wave calls visit(e, ev, own) with the edge index e[x] of the lane's
// position x, whether it is valid and its owner lane, so the caller can keep B
// loads in flight (and fetch per-vertex values of the owner with __shfl).
template <int B, class Visit>
__device__ __forceinline__ void wave_expand(uint32_t eb, uint32_t deg, Visit visit) {
  const uint32_t incl = wave_incl_scan(deg);
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  const uint32_t excl = incl - deg;
  for (uint32_t off = 0; off < total; off += B * 64) {
    uint32_t e[B], own[B];
    bool ev[B];
#pragma unroll
    for (int x = 0; x < B; ++x) {
      if (off + 64u * x >= total) {  // past the last edge for the whole wave
        e[x] = 0; own[x] = 0; ev[x] = false;
        continue;
      }
      const uint32_t q = off + 64 * x + lane_id();
      ev[x] = q < total;
      own[x] = wave_owner(incl, ev[x] ? q : total - 1);
      e[x] = __shfl(eb, (int)own[x]) + (q - __shfl(excl, (int)own[x]));
    }
    visit(e, ev, own);
  }
}

}  // namespace hs
