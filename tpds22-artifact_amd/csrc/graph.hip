// graph.hip -- persistent graph kernels on CSR uint32: single-source BFS,
// 64-source bit-parallel BFS with a push/pull schedule, near/far SSSP and
// deterministic pull PageRank.  One co-resident grid, grid barriers between
// phases (any hs_barrier_kind).
//
// Reference: LONESTAR/apps/{bfs,sssp,pr}/kernel_gsrb.cu (one persistent kernel
// per app, a barrier per round), LONESTAR/rt/include/worklist.h (frontier
// queues with warp-aggregated pushes) and LONESTAR/rt/include/internal.h (NP
// expansion: TB / warp / fine-grained).  Here the expansion is wave64-native:
// each wave takes up to 64 frontier vertices, scans their degrees with DPP and
// hands the concatenated edge list out 64 edges at a time (fine grained, the
// owner found by a 6-step shuffle search); vertices of degree >= 256 are
// expanded by the whole workgroup (BFS/SSSP) or split into fixed chunks that
// every wave of the grid shares (multi-source BFS, PageRank).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "hs_graph.h"
#include "hs_internal.hpp"
#include "hs_wave.hpp"
#include "hs_worklist.hpp"

using namespace hs;

namespace hs_host {
int prepare_barrier(hs_arena *ar, int kind, hipStream_t s);
int launch_persistent(const void *k, void **args, int kind, int num_wgs, int wg_size, size_t lds,
                      hipStream_t s);
}  // namespace hs_host

namespace {

constexpr int kWg = 256;
constexpr int32_t kInf = 0x7FFFFFFF;
constexpr uint32_t kHeavy = 256;     // degree at which a vertex leaves the fine-grained path
constexpr uint32_t kChunk = 4096;    // edges per shared chunk of a heavy vertex
constexpr int kCtr = 64;             // words between counters
// PageRank: 64-edge steps in flight per wave (light blocks) and per lane (heavy
// chunks).  The pull is bound by the L1's outstanding misses (the random
// gathers), not by how many a wave can issue: on R-MAT 22 an iteration takes
// 505 us with 8 and 8, 452 with 4 and 8, 439 with 4 and 2.  The step grouping
// does not change any sum's order.
constexpr int kPrSteps = 4;
constexpr int kPrHeavySteps = 2;

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * kWg + threadIdx.x; }
__device__ __forceinline__ uint32_t gthreads() { return gridDim.x * kWg; }
__device__ __forceinline__ uint32_t gwave() { return blockIdx.x * (kWg / 64) + (threadIdx.x >> 6); }
__device__ __forceinline__ uint32_t gwaves() { return gridDim.x * (kWg / 64); }

__device__ __forceinline__ int32_t ld_i32(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_u64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t lane) {
  const uint32_t lo = __shfl((uint32_t)v, (int)lane), hi = __shfl((uint32_t)(v >> 32), (int)lane);
  return ((uint64_t)hi << 32) | lo;
}

// ===========================================================================
// single-source BFS (LONESTAR/apps/bfs/kernel_gsrb.cu:400-626)
// Level-synchronous push.  A vertex reached at level L goes to one of two
// inputs of level L: the light queue (out-degree < kHeavy), expanded 256
// vertices per workgroup grab with the wave64 fine-grained scheme, or, when
// heavy, as chunks of kBfsChunk out-edges to the chunk list that every wave of
// the grid shares (so one hub does not serialize a level on one workgroup).
constexpr uint32_t kBfsChunk = 1024;

struct BfsArgs {
  uint32_t n;
  const uint32_t *row, *col;
  int32_t src;
  int32_t *level;
  uint32_t *q;    // [2][n] light frontier queues
  uint64_t *hq;   // [2][hcap] heavy chunk entries v | chunk << 32
  uint64_t *lq;   // [2][hcap] load-balanced mode: balanced entries v | first edge of v
                  // in the level's balanced edge space << 32
  uint64_t hcap;
  uint32_t *ctr;  // light counts [3], steal counters [3], chunk counts [3], balanced
                  // counts [3] (u64: entries << 32 | edges), kCtr words apart
  uint32_t *stats;
  uint32_t lb_limit;  // load-balanced mode: out-degree from which a vertex is heavy
  uint32_t ng, nx;
};

__device__ __forceinline__ uint32_t *cnt_at(uint32_t *c, uint32_t i) { return c + i * kCtr; }
__device__ __forceinline__ uint32_t *steal_at(uint32_t *c, uint32_t i) { return c + (3 + i) * kCtr; }
__device__ __forceinline__ uint32_t *hcnt_at(uint32_t *c, uint32_t i) { return c + (6 + i) * kCtr; }
__device__ __forceinline__ uint64_t *lbcnt_at(uint32_t *c, uint32_t i) {
  return reinterpret_cast<uint64_t *>(c + (9 + i) * kCtr);
}

__device__ __forceinline__ uint32_t bfs_chunks(uint32_t d) {
  return d >= kHeavy ? (d + kBfsChunk - 1) / kBfsChunk : 0u;
}

// visit (kernel_gsrb.cu:473-479): the reference writes level[dst]=LEVEL when it
// reads INF and may push duplicates; a CAS makes the push exactly-once.
// kBfsBatch candidates per lane at once: their level loads, CASes and degree
// loads are each in flight together before the wave-aggregated pushes.
constexpr int kBfsBatch = 8;
__device__ __forceinline__ void bfs_visit_batch(const BfsArgs &p, const uint32_t (&v)[kBfsBatch], int32_t L,
                                                uint32_t *qout, uint32_t *cout, uint64_t *hout,
                                                uint32_t *hcout) {
  int32_t lv[kBfsBatch];
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) lv[x] = v[x] != 0xFFFFFFFFu ? ld_i32(p.level + v[x]) : 0;
  bool won[kBfsBatch];
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) {
    won[x] = false;
    if (v[x] != 0xFFFFFFFFu && lv[x] == kInf) {
      int32_t expected = kInf;
      won[x] = __hip_atomic_compare_exchange_strong(p.level + v[x], &expected, L, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  uint32_t nc[kBfsBatch];
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) nc[x] = won[x] ? bfs_chunks(p.row[v[x] + 1] - p.row[v[x]]) : 0u;
  bool light[kBfsBatch];
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) light[x] = won[x] && nc[x] == 0;
  wave_push_batch<kBfsBatch>(qout, cout, light, v);
  wave_push_chunks_batch<kBfsBatch>(hout, hcout, nc, v);
}

// Load-balanced mode (the -l path of the reference: the inspector that moves
// vertices of out-degree >= DEGREE_LIMIT to a work list with their degrees, the
// prefix sum over those degrees and the kernel that gives every thread an equal
// share of their edges, LONESTAR/apps/bfs/kernel_gsrb.cu:343-398,569-579,
// LONESTAR/rt/include/thread_work.h:9-79).  Here the prefix sum is taken as the
// balanced vertices are discovered: one 64-bit fetch-and-add per wave on the
// counter (entries << 32 | their out-edges) returns both the wave's first list
// slot and its first edge, so the list is ordered by first edge; the next level
// splits the edges [0, total) evenly over every wave of the grid.  Vertices
// below the limit take the paths of the plain mode (chunk list from kHeavy,
// fine-grained below), as the reference's take its NP expansion.
__device__ __forceinline__ void bfs_visit_batch_lb(const BfsArgs &p, const uint32_t (&v)[kBfsBatch], int32_t L,
                                                   uint32_t *qout, uint32_t *cout, uint64_t *hout,
                                                   uint32_t *hcout, uint64_t *lout, uint64_t *lcout) {
  bool won[kBfsBatch];
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) {
    won[x] = false;
    if (v[x] != 0xFFFFFFFFu && ld_i32(p.level + v[x]) == kInf) {
      int32_t expected = kInf;
      won[x] = __hip_atomic_compare_exchange_strong(p.level + v[x], &expected, L, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  uint32_t d[kBfsBatch], nc[kBfsBatch];
  bool light[kBfsBatch];
  uint32_t hc = 0, hd = 0;
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) {
    d[x] = won[x] ? p.row[v[x] + 1] - p.row[v[x]] : 0u;
    const bool bal = won[x] && d[x] >= p.lb_limit;
    nc[x] = won[x] && !bal ? bfs_chunks(d[x]) : 0u;
    light[x] = won[x] && !bal && nc[x] == 0;
    if (!bal) d[x] = 0;
    hc += bal ? 1u : 0u;
    hd += d[x];
  }
  wave_push_batch<kBfsBatch>(qout, cout, light, v);
  wave_push_chunks_batch<kBfsBatch>(hout, hcout, nc, v);
  const uint32_t ic = wave_incl_scan(hc), id = wave_incl_scan(hd);
  const uint32_t tc = __builtin_amdgcn_readlane(ic, 63), td = __builtin_amdgcn_readlane(id, 63);
  if (tc == 0) return;
  uint64_t base = 0;
  if (lane_id() == 63)
    base = __hip_atomic_fetch_add(lcout, ((uint64_t)tc << 32) | td, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t b_lo = __builtin_amdgcn_readlane((uint32_t)base, 63);
  const uint32_t b_hi = __builtin_amdgcn_readlane((uint32_t)(base >> 32), 63);
  uint32_t slot = b_hi + ic - hc, eb = b_lo + id - hd;
#pragma unroll
  for (int x = 0; x < kBfsBatch; ++x) {
    if (d[x]) {
      lout[slot++] = v[x] | ((uint64_t)eb << 32);
      eb += d[x];
    }
  }
}

// the entry of the balanced list whose edge range holds edge q: the last entry
// whose first edge is <= q (entries ordered by first edge, the first at 0)
__device__ __forceinline__ uint32_t bfs_lb_owner(const uint64_t *lin, uint32_t nl, uint32_t q) {
  uint32_t lo = 0, hi = nl;  // first(lo) <= q < first(hi)
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint32_t)(lin[mid] >> 32) <= q) lo = mid; else hi = mid;
  }
  return lo;
}

template <bool LB>
__device__ __forceinline__ void bfs_visit(const BfsArgs &p, const uint32_t (&v)[kBfsBatch], int32_t L,
                                          uint32_t *qout, uint32_t *cout, uint64_t *hout, uint32_t *hcout,
                                          uint64_t *lout, uint64_t *lcout) {
  if constexpr (LB) bfs_visit_batch_lb(p, v, L, qout, cout, hout, hcout, lout, lcout);
  else bfs_visit_batch(p, v, L, qout, cout, hout, hcout);
}

template <int KIND, bool LB>
__device__ __forceinline__ void bfs_run(ArenaView a, const BfsArgs &p) {
  GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nx);
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) p.level[v] = (v == (uint32_t)p.src) ? 0 : kInf;
  if (gtid() == 0) {
    for (int i = 0; i < 3; ++i) {
      *cnt_at(p.ctr, i) = 0;
      *steal_at(p.ctr, i) = 0;
      *hcnt_at(p.ctr, i) = 0;
      if (LB) *lbcnt_at(p.ctr, i) = 0;
    }
    const uint32_t d = p.row[p.src + 1] - p.row[p.src];
    const uint32_t nc = bfs_chunks(d);
    if (LB && d >= p.lb_limit) {
      p.lq[0] = (uint32_t)p.src;
      *lbcnt_at(p.ctr, 0) = (1ull << 32) | d;
    } else if (nc) {
      for (uint32_t c = 0; c < nc; ++c) p.hq[c] = (uint32_t)p.src | ((uint64_t)c << 32);
      *hcnt_at(p.ctr, 0) = nc;
    } else {
      p.q[0] = (uint32_t)p.src;
      *cnt_at(p.ctr, 0) = 1;
    }
  }
  if (gb.sync<KIND>()) return;
  uint32_t k = 0;
  for (;; ++k) {
    const Frontier fr{p.q, p.ctr, p.ctr + 3 * kCtr, p.n, (uint32_t)kCtr};
    const uint32_t nin = fr.in_count(k);
    const uint32_t nh = ld_relaxed(hcnt_at(p.ctr, k % 3));
    uint32_t nl = 0, el = 0;  // load-balanced mode: balanced entries and their edges
    if constexpr (LB) {
      const uint64_t lc = __hip_atomic_load(lbcnt_at(p.ctr, k % 3), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      nl = (uint32_t)(lc >> 32);
      el = (uint32_t)lc;
    }
    if (nin == 0 && nh == 0 && nl == 0) break;
    const uint32_t *qin = fr.in(k);
    uint32_t *qout = fr.out(k);
    const uint64_t *hin = p.hq + (size_t)(k & 1) * p.hcap;
    uint64_t *hout = p.hq + (size_t)((k + 1) & 1) * p.hcap;
    const uint64_t *lin = LB ? p.lq + (size_t)(k & 1) * p.hcap : nullptr;
    uint64_t *lout = LB ? p.lq + (size_t)((k + 1) & 1) * p.hcap : nullptr;
    uint32_t *cout = fr.out_count(k);
    uint32_t *hcout = hcnt_at(p.ctr, (k + 1) % 3);
    uint64_t *lcout = lbcnt_at(p.ctr, (k + 1) % 3);
    if (gtid() == 0) {  // counters last used two levels ago
      fr.clear_for_next(k);
      st_relaxed(hcnt_at(p.ctr, (k + 2) % 3), 0u);
      if (LB) __hip_atomic_store(lbcnt_at(p.ctr, (k + 2) % 3), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int32_t L = (int32_t)k + 1;
    if constexpr (LB) {
      // the balanced vertices' edges [0, el), an equal share per wave of the
      // grid; a lane's owner entry only moves forward as its edge does (an
      // entry holds >= kHeavy edges, a step of a lane 64 apart)
      const uint32_t W = gwaves(), w = gwave();
      const uint32_t q_lo = (uint32_t)((uint64_t)el * w / W), q_hi = (uint32_t)((uint64_t)el * (w + 1) / W);
      if (q_lo < q_hi) {
        uint32_t j = bfs_lb_owner(lin, nl, q_lo);
        uint64_t ent = lin[j];
        uint32_t rb = p.row[(uint32_t)ent];
        uint32_t next = j + 1 < nl ? (uint32_t)(lin[j + 1] >> 32) : 0xFFFFFFFFu;  // first edge of j + 1
        for (uint32_t q0 = q_lo; q0 < q_hi; q0 += kBfsBatch * 64) {
          uint32_t vb[kBfsBatch];
#pragma unroll
          for (int x = 0; x < kBfsBatch; ++x) {
            const uint32_t q = q0 + 64 * x + lane_id();
            vb[x] = 0xFFFFFFFFu;
            if (q < q_hi) {
              while (next <= q) {
                ++j;
                ent = lin[j];
                rb = p.row[(uint32_t)ent];
                next = j + 1 < nl ? (uint32_t)(lin[j + 1] >> 32) : 0xFFFFFFFFu;
              }
              vb[x] = p.col[rb + (q - (uint32_t)(ent >> 32))];
            }
          }
          bfs_visit<LB>(p, vb, L, qout, cout, hout, hcout, lout, lcout);
        }
      }
    }
    // heavy chunks, spread over every wave of the grid
    for (uint32_t i = gwave(); i < nh; i += gwaves()) {
      const uint64_t ent = hin[i];
      const uint32_t u = (uint32_t)ent, c = (uint32_t)(ent >> 32);
      const uint32_t eb = p.row[u] + c * kBfsChunk, ee = min(p.row[u + 1], eb + kBfsChunk);
      for (uint32_t e0 = eb; e0 < ee; e0 += kBfsBatch * 64) {
        uint32_t vb[kBfsBatch];
#pragma unroll
        for (int x = 0; x < kBfsBatch; ++x) {
          const uint32_t e = e0 + 64 * x + lane_id();
          vb[x] = e < ee ? p.col[e] : 0xFFFFFFFFu;
        }
        bfs_visit<LB>(p, vb, L, qout, cout, hout, hcout, lout, lcout);
      }
    }
    // light queue: 256 vertices per workgroup grab, wave64 fine-grained expansion
    __shared__ uint32_t s_base;
    for (;;) {
      if (threadIdx.x == 0) s_base = fr.grab(k, (uint32_t)kWg);
      __syncthreads();
      const uint32_t base = s_base;
      __syncthreads();
      if (base >= nin) break;
      const uint32_t i = base + threadIdx.x;
      const bool valid = i < nin;
      const uint32_t u = valid ? qin[i] : 0u;
      uint32_t b = 0, deg = 0;
      if (valid) { b = p.row[u]; deg = p.row[u + 1] - b; }
      wave_expand<kBfsBatch>(b, deg, [&](const uint32_t (&e)[kBfsBatch], const bool (&ev)[kBfsBatch],
                                         const uint32_t (&)[kBfsBatch]) {
        uint32_t vb[kBfsBatch];
#pragma unroll
        for (int x = 0; x < kBfsBatch; ++x) vb[x] = ev[x] ? p.col[e[x]] : 0xFFFFFFFFu;
        bfs_visit<LB>(p, vb, L, qout, cout, hout, hcout, lout, lcout);
      });
    }
    if (gb.sync<KIND>()) return;
  }
  if (gtid() == 0 && p.stats) p.stats[0] = k;
}

template <int KIND>
__global__ void __launch_bounds__(kWg) __attribute__((amdgpu_waves_per_eu(8))) k_bfs(ArenaView a, BfsArgs p) {
  bfs_run<KIND, false>(a, p);
}

template <int KIND>
__global__ void __launch_bounds__(kWg) __attribute__((amdgpu_waves_per_eu(8))) k_bfs_lb(ArenaView a, BfsArgs p) {
  bfs_run<KIND, true>(a, p);
}

// ===========================================================================
// multi-source BFS: bit s of fr[v] / vis[v] / nx[v] belongs to source s
struct MsArgs {
  uint32_t n;
  uint64_t m;
  const uint32_t *row, *col, *rrow, *rcol;
  const int32_t *srcs;
  int nsrc;
  int32_t *levels;          // [nsrc][n]
  uint64_t *fb[2];          // [n] each: level L's frontier is fb[L & 1], its next frontier fb[(L + 1) & 1]
  uint64_t *vis;            // [n]
  uint64_t *fbm[2];         // [ceil(n / 64)] each: bit v & 63 of word v >> 6 of fbm[L & 1] is
                            // "v is in level L's frontier"
  uint64_t *hlist;          // [2 lists][2 parities][hcap] entries v | chunk << 32
  uint32_t *hcnt;           // [2 lists][2 parities] counters, kCtr apart
  uint64_t hcap;
  uint64_t *part;           // [2][W][4] per-WG stats
  uint64_t *planes;         // [kMsPlanes][n]: bit s of plane k = bit k of (level + 1) of (s, v)
  int schedule;
  const float *mlp;
  const uint8_t *forced;    // [HS_MSBFS_MAX_LEVELS] (device)
  uint32_t *trace;
  uint32_t ng, nxcd;
};

// Discovery levels are kept as bit-planes while the BFS runs: a vertex reached
// by source s at level L < 255 sets bit s of planes[k][v] for every bit k of
// L + 1 (a read-modify-write of 8-byte words by the vertex's owner thread,
// coalesced), and one final pass decodes them and writes levels[s][v] once per
// pair, a full row of 64 vertices per store instruction.  Writing levels[s][v]
// at discovery instead hits every cache line of the 64 x n array at several
// levels with partial writes.  Levels >= 255 (deep graphs) are written directly.
// Plane k > 0 is written whole by the update pass of the first level with bit k
// (L + 1 = 2^k) and the decode reads only the planes written: a BFS of 8 levels
// clears and reads 4 of the 8.
constexpr int kMsPlanes = 8;
// edges per chunk of a heavy vertex in the MS-BFS lists: at R-MAT scale 22 the
// push from the level-2 hubs takes 127 us with 2048, 153 us with 4096
constexpr uint32_t kMsChunk = 2048;

struct MsStats {
  uint64_t nf, mf, mu, disc;
};

constexpr int kPush = 0, kPull = 1;
constexpr int kMlpIn = 8, kMlpH = 64, kMlpOut = 8;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + 2 * (kMlpH * kMlpH + kMlpH) + kMlpOut * kMlpH + kMlpOut;
// A level's action: bit 0 direction (push / pull), bit 1 expansion on half of
// the workgroups (the others go straight to the barrier), bit 2 long poll
// backoff in the barrier that ends the expansion.
constexpr uint32_t kActHalf = 2, kActLongBackoff = 4;
constexpr uint32_t kLongBackoff = 32;  // s_sleep units of a long-backoff level's polls

__device__ __forceinline__ uint64_t *hl_at(const MsArgs &p, int list, uint32_t parity) {
  return p.hlist + (size_t)(list * 2 + parity) * p.hcap;
}
// The frontier and the next frontier swap buffers every level: the update pass
// after level L leaves the new frontier where the expansion wrote it and clears
// only the words of the old frontier (found from its bitmap), instead of
// copying every word and clearing the other buffer.
__device__ __forceinline__ uint64_t *fr_at(const MsArgs &p, uint32_t L) { return p.fb[L & 1u]; }
__device__ __forceinline__ uint64_t *nx_at(const MsArgs &p, uint32_t L) { return p.fb[(L + 1u) & 1u]; }
__device__ __forceinline__ uint64_t *fbm_at(const MsArgs &p, uint32_t L) { return p.fbm[L & 1u]; }

__device__ __forceinline__ uint32_t *hc_at(const MsArgs &p, int list, uint32_t parity) {
  return p.hcnt + (list * 2 + parity) * kCtr;
}
// the pull's work is claimed from kMsGroups counters per level parity (after the
// four list counters)
constexpr uint32_t kMsGroups = 64;
constexpr uint32_t kMsCounters = 4 + 2 * kMsGroups;
__device__ __forceinline__ uint32_t *claim_at(const MsArgs &p, uint32_t parity, uint32_t g) {
  return p.hcnt + (4 + parity * kMsGroups + g) * kCtr;
}

// features of a level (all schedule-invariant; hsync.sched.features computes the
// same): frontier vertices, frontier out-edges, unvisited vertices' in-edges and
// the previous level's discoveries relative to the graph, its mean degree, the
// level, the out/in edge balance, and the graph size (costs change when the
// graph stops fitting the caches)
__device__ __forceinline__ float lg1(float x) { return log2f(1.f + x); }
__device__ __forceinline__ void ms_features(const MsStats &s, uint32_t n, uint64_t m, uint32_t L,
                                            float *f) {
  const float ln = lg1((float)n), lm = lg1((float)m);
  f[0] = (lg1((float)s.nf) - ln) / 32.f;
  f[1] = (lg1((float)s.mf) - lm) / 32.f;
  f[2] = (lg1((float)s.mu) - lm) / 32.f;
  f[3] = (lg1((float)s.disc) - ln) / 32.f;
  f[4] = (lm - ln) / 8.f;
  f[5] = (float)L / 32.f;
  f[6] = (lg1((float)s.mf) - lg1((float)s.mu)) / 8.f;
  f[7] = ln / 32.f;
}

// 8 -> 64 -> 64 -> 64 -> 8 MLP, ReLU; wave 0 of the workgroup evaluates it with
// one hidden unit per lane.  Layout: W1[64][8] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[8][64] b4[8], then one more float: the judge margin.
// The plain schedule judges: the network's best action replaces the plain
// action `plain` only when its logit leads the plain action's by at least the
// margin of the graph's size band (a margin of +inf keeps the plain schedule).
// Returns the action.
__device__ int mlp_decide(const float *w, const float *x, int plain) {
  const uint32_t j = lane_id();
  const float *W1 = w, *b1 = W1 + kMlpH * kMlpIn, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH,
              *W3 = b2 + kMlpH, *b3 = W3 + kMlpH * kMlpH, *W4 = b3 + kMlpH, *b4 = W4 + kMlpOut * kMlpH;
  // the judge's margin for this graph's size band: w[P] up to log2(n + 1) =
  // w[P + 2] (feature 7 is log2(n + 1) / 32), w[P + 1] above it
  const float margin = x[7] * 32.f <= w[kMlpParams + 2] ? w[kMlpParams] : w[kMlpParams + 1];
  // no lead reaches +inf: the plain schedule, without evaluating the network
  // (every workgroup evaluates it per level, 37 KiB of weights each)
  if (__builtin_isinf(margin) && margin > 0.f) return plain;
  float h = b1[j];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) h = fmaf(W1[j * kMlpIn + i], x[i], h);
  h = fmaxf(h, 0.f);
  float h2 = b2[j];
  for (int k = 0; k < kMlpH; ++k) h2 = fmaf(W2[j * kMlpH + k], __shfl(h, k), h2);
  h2 = fmaxf(h2, 0.f);
  float h3 = b3[j];
  for (int k = 0; k < kMlpH; ++k) h3 = fmaf(W3[j * kMlpH + k], __shfl(h2, k), h3);
  h3 = fmaxf(h3, 0.f);
  int best = 0;
  float ob = 0.f, op = 0.f;
  for (int q = 0; q < kMlpOut; ++q) {
    float o = W4[q * kMlpH + j] * h3;
    for (int k = 32; k >= 1; k >>= 1) o += __shfl_xor(o, k);  // fixed butterfly order: identical in every wave
    o += b4[q];
    if (q == 0 || o > ob) { ob = o; best = q; }
    if (q == plain) op = o;
  }
  return (ob - op >= margin) ? best : plain;
}

// the plain schedule: pull when the frontier's out-edges exceed half of the
// unvisited vertices' in-edges (the pull scans in-edges, the push out-edges)
__device__ __forceinline__ int heuristic_dir(const MsStats &s) { return s.mf * 2 > s.mu ? kPull : kPush; }

// Blocks of 64 vertices are visited in a scattered order: block j of a wave's
// share is scatter_block(j) (a bijection of [0, 2^k): xor-shifts and odd
// multiplies, so that neighbouring j land far apart; values >= nblk are
// skipped).  A plain stride of gwaves() blocks would hand every wave blocks with
// the same low index bits, and R-MAT degrees follow the bits of the vertex id.
__device__ __forceinline__ uint32_t scatter_block(uint32_t j, uint32_t k) {
  const uint32_t mask = k >= 32 ? 0xFFFFFFFFu : ((1u << k) - 1u);
  const uint32_t sh = (k + 1) / 2;
  uint32_t h = j;
  h ^= h >> sh;
  h = (h * 0x9E3779B1u) & mask;
  h ^= h >> sh;
  h = (h * 0x85EBCA6Bu) & mask;
  h ^= h >> sh;
  return h;
}

__device__ __forceinline__ uint32_t log2_ceil(uint32_t x) {
  uint32_t k = 0;
  while ((1ull << k) < x) ++k;
  return k;
}

// The chunk entries of an update pass's heavy vertices are staged in LDS and
// published with one counter atomic per list at the end of the pass; a wave
// takes the slots of all its lanes' entries with one atomic.  A counter atomic
// per heavy vertex serialises the workgroups that hold many of them: R-MAT
// hubs have ids with few set bits, so the workgroup of ids 0..255 mod the
// grid's stride holds ~10% of all heavy vertices (R-MAT scale 22, edge factor
// 16: 2380 of ~25k vertices of in-degree >= 256) and took twice the pass's
// median time.  Entries that do not fit into the stage go straight to the
// global list, again one atomic per wave.
constexpr uint32_t kMsStage = 512;  // staged entries per list and workgroup

struct MsStage {
  uint64_t ent[2][kMsStage];
  uint32_t cnt[2];   // slots handed out (may pass kMsStage)
  uint32_t fill[2];  // end of the filled prefix of ent
  uint32_t base[2];
};

// adds chunks 0..nc-1 of this lane's vertex v to list l (nc = 0: nothing);
// every lane of the wave calls it
__device__ __forceinline__ void ms_stage_add(MsStage &s, int l, uint32_t v, uint32_t nc, uint64_t *gl,
                                             uint32_t *gc, uint64_t cap) {
  if (!__ballot(nc != 0)) return;
  const uint32_t lane = lane_id();
  uint32_t inc = nc;  // inclusive prefix sum of nc over the wave
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc += y;
  }
  const uint32_t total = __shfl(inc, 63), ex = inc - nc;
  uint32_t at = lane == 63 ? atomicAdd(&s.cnt[l], total) : 0u;
  at = __shfl(at, 63);
  if (at + total <= kMsStage) {
    // slots are handed out in order, so the ranges that fit form a prefix
    for (uint32_t c = 0; c < nc; ++c) s.ent[l][at + ex + c] = v | ((uint64_t)c << 32);
    if (lane == 63) atomicMax(&s.fill[l], at + total);
  } else {
    uint32_t g = lane == 63 ? add_relaxed(gc, total) : 0u;
    g = __shfl(g, 63);
    for (uint32_t c = 0; c < nc && (uint64_t)g + ex + c < cap; ++c) gl[g + ex + c] = v | ((uint64_t)c << 32);
  }
}

// publishes the staged entries (every thread of the workgroup calls it)
__device__ void ms_stage_flush(MsStage &s, uint64_t *gl0, uint64_t *gl1, uint32_t *gc0, uint32_t *gc1,
                               uint64_t cap) {
  __syncthreads();
  if (threadIdx.x < 2) {
    const uint32_t f = s.fill[threadIdx.x];
    s.base[threadIdx.x] = f ? add_relaxed(threadIdx.x ? gc1 : gc0, f) : 0u;
  }
  __syncthreads();
  for (int l = 0; l < 2; ++l) {
    uint64_t *gl = l ? gl1 : gl0;
    const uint32_t f = s.fill[l], b = s.base[l];
    for (uint32_t i = threadIdx.x; i < f; i += kWg)
      if ((uint64_t)b + i < cap) gl[b + i] = s.ent[l][i];
  }
  __syncthreads();
  if (threadIdx.x < 2) { s.cnt[threadIdx.x] = 0; s.fill[threadIdx.x] = 0; }
  __syncthreads();
}

// one pass over all vertices: fold next into frontier/visited (unless init),
// write discovery levels, accumulate the next level's statistics and its
// heavy-vertex chunk lists
__device__ void ms_update(const MsArgs &p, uint32_t L, bool init, uint64_t allmask) {
  MsStats st{0, 0, 0, 0};
  const uint32_t lp = (L + 1) & 1u;  // lists for level L + 1 are stored at parity (L + 1) & 1
  uint64_t *hpush = hl_at(p, 0, lp), *hpull = hl_at(p, 1, lp);
  uint32_t *cpush = hc_at(p, 0, lp), *cpull = hc_at(p, 1, lp);
  __shared__ MsStage stg;
  if (threadIdx.x < 2) { stg.cnt[threadIdx.x] = 0; stg.fill[threadIdx.x] = 0; }
  __syncthreads();
  // blocks of kWg vertices in the scattered order (the workgroup of a plain
  // stride would hold every block whose middle id bits are zero: the hubs)
  const uint32_t nblk = (p.n + kWg - 1) / kWg, kb = log2_ceil(nblk), nj = 1u << kb;
  // one vertex: vis, the old frontier's clearing, the new bitmap, the levels,
  // the statistics and the heavy lists (every thread of the workgroup calls it)
  auto vertex = [&](uint32_t v, uint64_t nxv, uint64_t visv, uint64_t old) {
    const bool valid = v < p.n;
    const uint32_t w0 = v & ~63u;  // this wave's first vertex
    if (!init && nxv) {
      visv |= nxv;
      p.vis[v] = visv;
    }
    if (!init && ((old >> lane_id()) & 1u)) fr_at(p, L)[v] = 0;  // level L's frontier becomes L + 1's next
    {  // frontier bitmap of level L + 1, one word per wave (64 consecutive vertices)
      const uint64_t fm = __ballot(nxv != 0);
      if (lane_id() == 0 && w0 < p.n) fbm_at(p, L + 1)[w0 >> 6] = fm;
    }
    if (!init && L + 1 < (1u << kMsPlanes)) {
      const uint32_t l1 = L + 1;
      if ((l1 & (l1 - 1u)) == 0) {
        // L + 1 = 2^k: the first level with bit k, and this pass visits every
        // vertex, so plane k is written here instead of cleared up front
        if (valid) p.planes[(size_t)__builtin_ctz(l1) * p.n + v] = nxv;
      } else if (nxv) {  // bit-planes of L + 1
        for (int k = 0; k < kMsPlanes; ++k)
          if (l1 >> k & 1) p.planes[(size_t)k * p.n + v] |= nxv;
      }
    } else if (!init) {
      // deep levels: one source at a time for the whole wave, so each store
      // instruction covers up to 64 consecutive vertices of one row
      uint64_t any = nxv;
      for (int k = 32; k >= 1; k >>= 1) any |= shfl64(any, lane_id() ^ k);
      while (any) {
        const int src = __builtin_ctzll(any);
        any &= any - 1;
        if (nxv >> src & 1) p.levels[(size_t)src * p.n + v] = (int32_t)L;
      }
    }
    const bool active = nxv != 0;
    const bool open = valid && visv != allmask;
    // degrees only where the statistics use them: out-degrees of the frontier,
    // in-degrees of the vertices still missing a source
    const uint32_t od = active ? p.row[v + 1] - p.row[v] : 0u;
    const uint32_t id = open ? p.rrow[v + 1] - p.rrow[v] : 0u;
    st.nf += active;
    st.mf += active ? od : 0u;
    st.mu += open ? id : 0u;
    st.disc += (uint64_t)__builtin_popcountll(nxv);
    ms_stage_add(stg, 0, v, active && od >= kHeavy ? (od + kMsChunk - 1) / kMsChunk : 0u, hpush, cpush, p.hcap);
    ms_stage_add(stg, 1, v, open && id >= kHeavy ? (id + kMsChunk - 1) / kMsChunk : 0u, hpull, cpull, p.hcap);
  };
  // two blocks per round, both blocks' loads before either's stores: the pass
  // is bound by the latency of these loads, not by their bytes
  for (uint32_t j = blockIdx.x; j < nj; j += 2 * gridDim.x) {
    uint32_t vv[2];
    uint64_t nx2[2], vis2[2], old2[2];
    bool use[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t jj = j + u * gridDim.x;
      const uint32_t blk = jj < nj ? scatter_block(jj, kb) : nblk;
      use[u] = blk < nblk;  // the same for the whole workgroup
      vv[u] = blk * kWg + threadIdx.x;
      nx2[u] = vis2[u] = old2[u] = 0;
      if (use[u] && vv[u] < p.n) {
        nx2[u] = fr_at(p, L + 1)[vv[u]];  // level L + 1's frontier, in place (the expansions add only bits not in vis)
        vis2[u] = p.vis[vv[u]];
      }
      if (use[u] && !init && (vv[u] & ~63u) < p.n) old2[u] = fbm_at(p, L)[vv[u] >> 6];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (use[u]) vertex(vv[u], nx2[u], vis2[u], old2[u]);
  }
  ms_stage_flush(stg, hpush, hpull, cpush, cpull, p.hcap);
  // workgroup totals (fixed order) -> part[(L+1)&1][wg]
  __shared__ uint64_t red[4][kWg / 64];
  uint64_t vals[4] = {st.nf, st.mf, st.mu, st.disc};
  for (int q = 0; q < 4; ++q) {
    uint64_t x = vals[q];
    for (int k = 32; k >= 1; k >>= 1) x += shfl64(x, lane_id() ^ k);
    if (lane_id() == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    uint64_t t = 0;
    for (int w = 0; w < kWg / 64; ++w) t += red[threadIdx.x][w];
    p.part[((size_t)lp * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = t;
  }
  __syncthreads();
}

__device__ MsStats ms_totals(const MsArgs &p, uint32_t parity) {
  __shared__ uint64_t red[4][kWg / 64];
  uint64_t acc[4] = {0, 0, 0, 0};
  for (uint32_t w = threadIdx.x; w < gridDim.x; w += kWg)
    for (int q = 0; q < 4; ++q) acc[q] += p.part[((size_t)parity * gridDim.x + w) * 4 + q];
  for (int q = 0; q < 4; ++q) {
    uint64_t x = acc[q];
    for (int k = 32; k >= 1; k >>= 1) x += shfl64(x, lane_id() ^ k);
    if (lane_id() == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  MsStats s;
  uint64_t t[4];
  for (int q = 0; q < 4; ++q) {
    t[q] = 0;
    for (int w = 0; w < kWg / 64; ++w) t[q] += red[q][w];
  }
  __syncthreads();
  s.nf = t[0]; s.mf = t[1]; s.mu = t[2]; s.disc = t[3];
  return s;
}

constexpr int kMsBatch = 4;  // 64-edge steps in flight per wave

// push: out-edges of frontier vertices, atomicOr of the new bits into nx
__device__ void ms_push(const MsArgs &p, uint32_t L, uint32_t nact) {
  if (gwave() >= nact) return;  // this level expands on the first nact waves only
  const uint64_t *fr = fr_at(p, L);
  uint64_t *nx = nx_at(p, L);
  const uint32_t lane = lane_id();
  const uint32_t nblk = (p.n + 63) / 64, kb = log2_ceil(nblk);
  for (uint32_t j = gwave(); j < (1u << kb); j += nact) {
    const uint32_t blk = scatter_block(j, kb);
    if (blk >= nblk) continue;
    const uint32_t v = blk * 64 + lane;
    const uint64_t f = v < p.n ? fr[v] : 0ull;
    uint32_t b = 0, deg = 0;
    if (f) { b = p.row[v]; deg = p.row[v + 1] - b; }
    wave_expand<kMsBatch>(b, deg < kHeavy ? deg : 0u, [&](const uint32_t (&e)[kMsBatch],
                                                         const bool (&ev)[kMsBatch],
                                                         const uint32_t (&own)[kMsBatch]) {
      uint32_t dst[kMsBatch];
      uint64_t of[kMsBatch], vis[kMsBatch], nxv[kMsBatch];
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) {
        of[x] = shfl64(f, own[x]);
        dst[x] = ev[x] ? __builtin_nontemporal_load(p.col + e[x]) : 0u;
      }
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) vis[x] = ev[x] ? ld_u64(p.vis + dst[x]) : ~0ull;
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) nxv[x] = (of[x] & ~vis[x]) ? ld_u64(nx + dst[x]) : ~0ull;
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) {
        const uint64_t bits = of[x] & ~vis[x];
        if (bits && (nxv[x] & bits) != bits)
          __hip_atomic_fetch_or(nx + dst[x], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    });
  }
  (void)lane;
  // heavy vertices: chunks of kChunk out-edges, shared round-robin by all waves
  const uint32_t cnt = ld_relaxed(hc_at(p, 0, L & 1u));
  const uint64_t *hl = hl_at(p, 0, L & 1u);
  for (uint32_t i = gwave(); i < cnt && i < p.hcap; i += nact) {
    const uint64_t ent = hl[i];
    const uint32_t v = (uint32_t)ent, c = (uint32_t)(ent >> 32);
    const uint64_t f = fr[v];
    const uint32_t eb = p.row[v] + c * kMsChunk;
    const uint32_t ee = min(p.row[v + 1], eb + kMsChunk);
    for (uint32_t e0 = eb + lane_id(); e0 < ee; e0 += kMsBatch * 64) {
      uint32_t dst[kMsBatch];
      uint64_t vis[kMsBatch];
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x)
        dst[x] = e0 + 64 * x < ee ? __builtin_nontemporal_load(p.col + e0 + 64 * x) : 0xFFFFFFFFu;
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) vis[x] = dst[x] != 0xFFFFFFFFu ? ld_u64(p.vis + dst[x]) : ~0ull;
#pragma unroll
      for (int x = 0; x < kMsBatch; ++x) {
        const uint64_t bits = f & ~vis[x];
        if (bits && (ld_u64(nx + dst[x]) & bits) != bits)
          __hip_atomic_fetch_or(nx + dst[x], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

__device__ __forceinline__ uint64_t lds_ld64(const uint64_t *a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
  for (int k = 32; k >= 1; k >>= 1) x |= shfl64(x, lane_id() ^ k);
  return x;
}
// frontier bits of in-neighbour u: the bitmap word (a few hundred KiB, kept
// in L2) first, fr[u] only for frontier vertices
__device__ __forceinline__ bool ms_in_frontier(uint64_t word, uint32_t u) {
  return (word >> (u & 63u)) & 1u;
}

// pull: in-edges of not-yet-complete vertices; light vertices accumulate in LDS
// and are stored by their owner (no atomics), heavy ones use atomicOr per chunk.
// In-neighbours are tested against the frontier bitmap before fr is read.
//
// Work is claimed, not dealt: the blocks of 64 vertices (scattered order) and
// then the heavy chunks form one sequence per group of workgroups, and each
// wave claims its next item from its group's counter while it works on the
// current one.  A fixed deal of the same blocks left the slowest wave at about
// twice the mean (R-MAT scale 22: in-edge counts per block are heavy-tailed).
// A light vertex's in-edges are read in two rounds.  The first step of the
// wave (kMsBatch x 64 positions) takes the first kMsFirst in-edges of every
// lane's vertex, then as many further edges as fit; the rest of the edges of
// the vertices still missing a source follow, packed again.  A vertex whose
// first edges found every source it was missing reads no more: at the level
// that reaches most of an R-MAT graph, ~80% of the light in-edges.  The number
// of steps stays that of one packed sequence when nothing is found early.
constexpr uint32_t kMsFirst = 4;
static_assert(kMsFirst <= (uint32_t)kMsBatch, "segment 1 must fit into the first step");

// Edge arrays are read with streaming loads (col in the push, rcol in the
// pull): each is read once per level, and cached they evict the vis, frontier
// and plane words that the update pass after the expansion reads again (R-MAT
// 22: the update passes after the big pulls 60 -> 51 us and 82 -> 62 us).
__device__ __forceinline__ void ms_pull_gather(const MsArgs &p, uint32_t L, const uint32_t (&e)[kMsBatch],
                                               const bool (&ev)[kMsBatch], const uint32_t (&own)[kMsBatch],
                                               uint64_t unv, uint64_t *acc) {
  const uint64_t *fr = fr_at(p, L), *fbm = fbm_at(p, L);
  uint32_t src[kMsBatch];
  uint64_t fm[kMsBatch], fb[kMsBatch];
#pragma unroll
  for (int x = 0; x < kMsBatch; ++x) src[x] = ev[x] ? __builtin_nontemporal_load(p.rcol + e[x]) : 0u;
#pragma unroll
  for (int x = 0; x < kMsBatch; ++x) fm[x] = ev[x] ? fbm[src[x] >> 6] : 0ull;
#pragma unroll
  for (int x = 0; x < kMsBatch; ++x) fb[x] = ms_in_frontier(fm[x], src[x]) ? ld_u64(fr + src[x]) : 0ull;
#pragma unroll
  for (int x = 0; x < kMsBatch; ++x) {
    const uint64_t bits = fb[x] & shfl64(unv, own[x]);
    if (bits) __hip_atomic_fetch_or(acc + own[x], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
}

__device__ __forceinline__ void ms_pull_light(const MsArgs &p, uint32_t L, uint64_t allmask, uint32_t blk,
                                              uint64_t *acc) {
  const uint32_t lane = lane_id();
  const uint32_t v = blk * 64 + lane;
  const uint64_t unv = v < p.n ? (~p.vis[v] & allmask) : 0ull;
  uint32_t b = 0, deg = 0;
  if (unv) { b = p.rrow[v]; deg = p.rrow[v + 1] - b; }
  const uint32_t ldeg = deg < kHeavy ? deg : 0u;
  // first round: segment 1 = the first kMsFirst edges of each lane, segment 2 =
  // the others, one sequence, its first kMsBatch x 64 positions
  const uint32_t d1 = ldeg < kMsFirst ? ldeg : kMsFirst, d2 = ldeg - d1;
  const uint32_t i1 = wave_incl_scan(d1), i2 = wave_incl_scan(d2);
  const uint32_t s1 = __builtin_amdgcn_readlane(i1, 63), s2 = __builtin_amdgcn_readlane(i2, 63);
  const uint32_t x1 = i1 - d1, x2 = i2 - d2;
  if (s1 + s2 == 0) return;
  {
    uint32_t e[kMsBatch], own[kMsBatch];
    bool ev[kMsBatch];
#pragma unroll
    for (int x = 0; x < kMsBatch; ++x) {
      if (64u * x >= s1 + s2) {  // past the last edge for the whole wave
        own[x] = 0; ev[x] = false; e[x] = 0;
        continue;
      }
      const uint32_t q = 64 * x + lane;
      const bool seg1 = q < s1;
      const uint32_t q2 = seg1 ? 0u : q - s1;
      const uint32_t o1 = wave_owner(i1, seg1 ? q : s1 - 1);  // s1 > 0 whenever any edge exists
      const uint32_t o2 = wave_owner(i2, q2 < s2 ? q2 : (s2 ? s2 - 1 : 0u));
      const uint32_t bo1 = __shfl(b, (int)o1), xo1 = __shfl(x1, (int)o1);
      const uint32_t bo2 = __shfl(b, (int)o2), xo2 = __shfl(x2, (int)o2);
      own[x] = seg1 ? o1 : o2;
      ev[x] = q < s1 + s2;
      e[x] = seg1 ? bo1 + (q - xo1) : bo2 + kMsFirst + (q2 - xo2);
    }
    ms_pull_gather(p, L, e, ev, own, unv, acc);
  }
  // second round: the segment-2 edges past the first round's positions, of the
  // vertices still missing a source
  const uint32_t used2 = kMsBatch * 64 - s1;  // segment-2 positions taken by the first round
  const uint32_t rest = i2 > used2 ? i2 - (x2 > used2 ? x2 : used2) : 0u;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // the lanes' LDS ORs before the reads
  const bool done = (lds_ld64(acc + lane) & unv) == unv;
  const uint32_t r = done ? 0u : rest;
  wave_expand<kMsBatch>(b + kMsFirst + (d2 - rest), r, [&](const uint32_t (&e)[kMsBatch],
                                                           const bool (&ev)[kMsBatch],
                                                           const uint32_t (&own)[kMsBatch]) {
    ms_pull_gather(p, L, e, ev, own, unv, acc);
  });
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (ldeg) {
    const uint64_t got = lds_ld64(acc + lane);
    if (got) nx_at(p, L)[v] = got;
  }
  acc[lane] = 0;
}

__device__ __forceinline__ void ms_pull_heavy(const MsArgs &p, uint32_t L, uint64_t allmask, uint64_t ent) {
  const uint64_t *fr = fr_at(p, L), *fbm = fbm_at(p, L);
  const uint32_t lane = lane_id();
  const uint32_t v = (uint32_t)ent, c = (uint32_t)(ent >> 32);
  const uint64_t unv = ~p.vis[v] & allmask;
  const uint32_t eb = p.rrow[v] + c * kMsChunk;
  const uint32_t ee = min(p.rrow[v + 1], eb + kMsChunk);
  uint64_t a = 0;
  for (uint32_t e00 = eb; e00 < ee; e00 += 8 * 64) {
    uint32_t src[8];
    uint64_t fm[8];
#pragma unroll
    for (int x = 0; x < 8; ++x)
      src[x] = e00 + 64 * x + lane < ee ? __builtin_nontemporal_load(p.rcol + e00 + 64 * x + lane) : 0xFFFFFFFFu;
#pragma unroll
    for (int x = 0; x < 8; ++x) fm[x] = src[x] != 0xFFFFFFFFu ? fbm[src[x] >> 6] : 0ull;
#pragma unroll
    for (int x = 0; x < 8; ++x) a |= ms_in_frontier(fm[x], src[x]) ? ld_u64(fr + src[x]) : 0ull;
    if ((wave_or64(a) & unv) == unv) break;  // this chunk found every missing source
  }
  a = wave_or64(a) & unv;
  if (lane == 0 && a) __hip_atomic_fetch_or(nx_at(p, L) + v, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void ms_pull(const MsArgs &p, uint32_t L, uint64_t allmask, uint32_t nact) {
  if (gwave() >= nact) return;
  __shared__ uint64_t acc_s[kWg];
  const uint32_t lane = lane_id();
  uint64_t *acc = acc_s + (threadIdx.x & ~63u);
  acc[lane] = 0;
  const uint32_t nblk = (p.n + 63) / 64, kb = log2_ceil(nblk);
  // groups: workgroup w of the active ones claims from counter w mod G, and
  // every counter has at least one active workgroup
  const uint32_t nwg = (nact + kWg / 64 - 1) / (kWg / 64);
  const uint32_t G = nwg < kMsGroups ? nwg : kMsGroups, g = blockIdx.x % G;
  const uint32_t nj = 1u << kb;
  const uint32_t nlight = g < nj ? (nj - g + G - 1) / G : 0u;  // j = g + t G < 2^kb
  uint32_t cnt = ld_relaxed(hc_at(p, 1, L & 1u));
  cnt = cnt < p.hcap ? cnt : (uint32_t)p.hcap;
  const uint32_t nheavy = g < cnt ? (cnt - g + G - 1) / G : 0u;
  const uint64_t *hl = hl_at(p, 1, L & 1u);
  uint32_t *ctr = claim_at(p, L & 1u, g);
  uint32_t t = __shfl(lane == 0 ? add_relaxed(ctr, 1u) : 0u, 0);
  while (t < nlight + nheavy) {
    const uint32_t next = __shfl(lane == 0 ? add_relaxed(ctr, 1u) : 0u, 0);
    if (t < nlight) {
      const uint32_t blk = scatter_block(g + t * G, kb);
      if (blk < nblk) ms_pull_light(p, L, allmask, blk, acc);
    } else {
      ms_pull_heavy(p, L, allmask, hl[g + (t - nlight) * G]);
    }
    t = next;
  }
}

template <int KIND>
__global__ void __launch_bounds__(kWg) k_msbfs(ArenaView a, MsArgs p) {
  GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nxcd);
  const uint64_t allmask = p.nsrc >= 64 ? ~0ull : ((1ull << p.nsrc) - 1ull);
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) { p.fb[0][v] = 0; p.fb[1][v] = 0; p.vis[v] = 0; }
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) p.planes[v] = 0;  // plane 0 (the others: ms_update)
  if (gtid() < kMsCounters) p.hcnt[gtid() * kCtr] = 0;
  if (gb.sync<KIND>()) return;
  if (blockIdx.x == 0 && threadIdx.x < (uint32_t)p.nsrc) {
    const uint32_t s = threadIdx.x, v = (uint32_t)p.srcs[s];
    __hip_atomic_fetch_or(fr_at(p, 1) + v, 1ull << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(p.vis + v, 1ull << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(p.planes + v, 1ull << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // L + 1 = 1
  }
  if (gb.sync<KIND>()) return;
  ms_update(p, 0, true, allmask);  // statistics + heavy lists of level 1
  if (gb.sync<KIND>()) return;
  __shared__ int s_dir;
  for (uint32_t L = 1;; ++L) {
    const MsStats st = ms_totals(p, L & 1u);
    if (st.nf == 0) {
      // every update is done (a barrier followed the last one): decode the
      // planes, one row segment of 64 vertices per store; unreached -> INF,
      // reached at a deep level -> already written
      // planes in use: 0 and every k with 2^k <= L (the updates reached L + 1 <= L)
      const int kuse = (32 - __builtin_clz(L)) < kMsPlanes ? (32 - __builtin_clz(L)) : kMsPlanes;
      for (uint32_t v = gtid(); v < p.n; v += gthreads()) {
        uint64_t pl[kMsPlanes];
#pragma unroll
        for (int k = 0; k < kMsPlanes; ++k) pl[k] = k < kuse ? p.planes[(size_t)k * p.n + v] : 0ull;
        const uint64_t vis = p.vis[v];
        for (int src = 0; src < p.nsrc; ++src) {
          uint32_t code = 0;
#pragma unroll
          for (int k = 0; k < kMsPlanes; ++k) code |= (uint32_t)((pl[k] >> src) & 1u) << k;
          // streaming stores: nothing here reads the 4 B x 64 x n levels again,
          // and cached they leave dirty lines that the next launch's first
          // passes pay for (clear and first update 81 -> 64 us at R-MAT 22)
          if (code) __builtin_nontemporal_store((int32_t)code - 1, p.levels + (size_t)src * p.n + v);
          else if (!((vis >> src) & 1u)) __builtin_nontemporal_store((int32_t)kInf, p.levels + (size_t)src * p.n + v);
        }
      }
      break;
    }
    if (threadIdx.x < 64) {
      int act;
      switch (p.schedule) {
        case 1: act = kPush; break;
        case 2: act = kPull; break;
        case 3: {
          float f[kMlpIn];
          ms_features(st, p.n, p.m, L, f);
          act = mlp_decide(p.mlp, f, heuristic_dir(st));
          break;
        }
        case 4: act = L <= HS_MSBFS_MAX_LEVELS ? (int)(p.forced[L - 1] & 7u) : kPush; break;
        default: act = heuristic_dir(st);
      }
      if (threadIdx.x == 0) s_dir = act;
    }
    __syncthreads();
    const int act = s_dir;
    const int dir = act & 1;
    const uint32_t nact = (act & kActHalf) ? ((gridDim.x + 1) / 2) * (kWg / 64) : gwaves();
    uint32_t t0 = 0, t1 = 0;
    if (p.trace && blockIdx.x == 0 && threadIdx.x == 0 && L <= HS_MSBFS_MAX_LEVELS) {
      const uint64_t c = wall_clock64();
      t0 = (uint32_t)c; t1 = (uint32_t)(c >> 32);
    }
    // counters of the lists for level L+1 were read at level L-1: clear them
    if (gtid() < 2) st_relaxed(hc_at(p, (int)gtid(), (L + 1) & 1u), 0u);
    if (gtid() < kMsGroups) st_relaxed(claim_at(p, (L + 1) & 1u, gtid()), 0u);
    if (dir == kPush) ms_push(p, L, nact);
    else ms_pull(p, L, allmask, nact);
    gb.poll_sleep = (act & kActLongBackoff) ? kLongBackoff : 0u;
    const int brc = gb.sync<KIND>();
    gb.poll_sleep = 0;
    if (brc) return;
    if (p.trace && blockIdx.x == 0 && threadIdx.x == 0 && L <= HS_MSBFS_MAX_LEVELS) {
      const uint64_t c = wall_clock64();
      uint32_t *tr = p.trace + (size_t)(L - 1) * HS_MSBFS_TRACE_WORDS;
      tr[0] = (uint32_t)act; tr[1] = (uint32_t)st.nf;
      tr[2] = (uint32_t)st.mf; tr[3] = (uint32_t)(st.mf >> 32);
      tr[4] = (uint32_t)st.mu; tr[5] = (uint32_t)(st.mu >> 32);
      tr[6] = (uint32_t)st.disc;
      tr[7] = t0; tr[8] = t1; tr[9] = (uint32_t)c; tr[10] = (uint32_t)(c >> 32);
    }
    ms_update(p, L, false, allmask);
    if (gb.sync<KIND>()) return;
  }
}

// ===========================================================================
// SSSP, near/far buckets (LONESTAR/apps/sssp/kernel_gsrb.cu:459-687).
struct SsspArgs {
  uint32_t n;
  const uint32_t *row, *col, *w;
  int32_t src, delta;
  int32_t *dist;
  uint32_t *nq;      // [2][n] near queues (light vertices)
  uint64_t *hq;      // [2][hcap] near chunk lists (heavy vertices, chunks of kBfsChunk edges)
  uint64_t hcap;
  uint32_t *fq;      // [3][n] far lists
  uint32_t *nstamp;  // [n] round stamp of the near list holding v
  uint32_t *fstamp;  // [n] epoch stamp of the far list holding v
  uint32_t *ctr;     // near counts [3] (0..2), steal [3] (3..5), far counts [3] (6..8),
                     // near chunk counts [3] (9..11)
  uint32_t *stats;
  uint32_t ng, nx;
};

constexpr int kSsBatch = 4;

// Entries of the next near input: light vertices to the queue, heavy ones as
// chunks to the chunk list (every wave of the grid shares them next round).
__device__ __forceinline__ void sssp_push_near(const SsspArgs &p, const bool (&near)[kSsBatch],
                                               const uint32_t (&v)[kSsBatch], uint32_t *qnear,
                                               uint32_t *cnear, uint64_t *hnear, uint32_t *hcnear) {
  uint32_t nc[kSsBatch];
#pragma unroll
  for (int x = 0; x < kSsBatch; ++x) nc[x] = near[x] ? bfs_chunks(p.row[v[x] + 1] - p.row[v[x]]) : 0u;
  bool light[kSsBatch];
#pragma unroll
  for (int x = 0; x < kSsBatch; ++x) light[x] = near[x] && nc[x] == 0;
  wave_push_batch<kSsBatch>(qnear, cnear, light, v);
  wave_push_chunks_batch<kSsBatch>(hnear, hcnear, nc, v);
}

// relaxation (kernel_gsrb.cu:415): atomicMin(dist[dst], dist[n] + w); the
// improved vertex joins the near list when within the current bucket bound,
// else the far list; the stamps keep each list free of duplicates.  A batch of
// kSsBatch candidates per lane (v == ~0 for none): their loads and atomics are
// each in flight together.
__device__ __forceinline__ void sssp_relax_batch(const SsspArgs &p, const uint32_t (&v)[kSsBatch],
                                                 const int32_t (&nd)[kSsBatch], int32_t bound,
                                                 uint32_t round_stamp, uint32_t *qnear, uint32_t *cnear,
                                                 uint64_t *hnear, uint32_t *hcnear, uint32_t epoch_stamp,
                                                 uint32_t *qfar, uint32_t *cfar) {
  int32_t cur[kSsBatch];
#pragma unroll
  for (int x = 0; x < kSsBatch; ++x) cur[x] = v[x] != 0xFFFFFFFFu ? ld_i32(p.dist + v[x]) : 0;
  bool better[kSsBatch];
#pragma unroll
  for (int x = 0; x < kSsBatch; ++x) {
    better[x] = false;
    if (v[x] != 0xFFFFFFFFu && nd[x] < cur[x]) {
      const int32_t old = __hip_atomic_fetch_min(p.dist + v[x], nd[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      better[x] = nd[x] < old;
    }
  }
  bool near[kSsBatch], far[kSsBatch];
#pragma unroll
  for (int x = 0; x < kSsBatch; ++x) {
    near[x] = far[x] = false;
    if (better[x]) {
      if (nd[x] <= bound)
        near[x] = __hip_atomic_exchange(p.nstamp + v[x], round_stamp, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) != round_stamp;
      else
        far[x] = __hip_atomic_exchange(p.fstamp + v[x], epoch_stamp, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT) != epoch_stamp;
    }
  }
  sssp_push_near(p, near, v, qnear, cnear, hnear, hcnear);
  wave_push_batch<kSsBatch>(qfar, cfar, far, v);
}

// Edge destinations and weights are read with streaming loads: each edge is
// read once per relaxation of its vertex, and cached they evict distances and
// stamps (R-MAT 22: 8.5 -> 8.0 ms per source; the same change slowed the
// 64-source SSSP by 2% and left BFS and PageRank unchanged, so only here).
template <int KIND>
__global__ void __launch_bounds__(kWg) __attribute__((amdgpu_waves_per_eu(8))) k_sssp(ArenaView a, SsspArgs p) {
  GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nx);
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) {
    p.dist[v] = (v == (uint32_t)p.src) ? 0 : kInf;
    p.nstamp[v] = 0;
    p.fstamp[v] = 0;
  }
  if (gtid() == 0) {
    for (int i = 0; i < 12; ++i) p.ctr[i * kCtr] = 0;
    const uint32_t nc = bfs_chunks(p.row[p.src + 1] - p.row[p.src]);
    if (nc) {
      for (uint32_t c = 0; c < nc; ++c) p.hq[c] = (uint32_t)p.src | ((uint64_t)c << 32);
      p.ctr[9 * kCtr] = nc;
    } else {
      p.nq[0] = (uint32_t)p.src;
      p.ctr[0] = 1;
    }
  }
  if (gb.sync<KIND>()) return;
  __shared__ uint32_t s_base;
  uint32_t r = 0, epoch = 0, relax_rounds = 0;
  for (;; ++r) {
    const uint32_t nin = ld_relaxed(p.ctr + (r % 3) * kCtr);
    const uint32_t nh = ld_relaxed(p.ctr + (9 + r % 3) * kCtr);
    uint32_t *qnear = p.nq + (size_t)((r + 1) & 1) * p.n;
    uint32_t *cnear = p.ctr + ((r + 1) % 3) * kCtr;
    uint64_t *hnear = p.hq + (size_t)((r + 1) & 1) * p.hcap;
    uint32_t *hcnear = p.ctr + (9 + (r + 1) % 3) * kCtr;
    if (gtid() == 0) {  // counters last used two rounds ago
      st_relaxed(p.ctr + ((r + 2) % 3) * kCtr, 0u);
      st_relaxed(p.ctr + (9 + (r + 2) % 3) * kCtr, 0u);
      st_relaxed(p.ctr + (3 + (r + 1) % 3) * kCtr, 0u);
    }
    const int32_t bound = (int32_t)min((int64_t)(epoch + 1) * p.delta, (int64_t)kInf - 1);
    if (nin > 0 || nh > 0) {
      uint32_t *qfar = p.fq + (size_t)(epoch % 3) * p.n;
      uint32_t *cfar = p.ctr + (6 + epoch % 3) * kCtr;
      // heavy chunks of the near input, spread over every wave of the grid
      const uint64_t *hin = p.hq + (size_t)(r & 1) * p.hcap;
      for (uint32_t i = gwave(); i < nh; i += gwaves()) {
        const uint64_t ent = hin[i];
        const uint32_t u = (uint32_t)ent, c = (uint32_t)(ent >> 32);
        const uint32_t eb = p.row[u] + c * kBfsChunk, ee = min(p.row[u + 1], eb + kBfsChunk);
        const int32_t du = ld_i32(p.dist + u);
        for (uint32_t e0 = eb; e0 < ee; e0 += kSsBatch * 64) {
          uint32_t vb[kSsBatch];
          int32_t nd[kSsBatch];
#pragma unroll
          for (int x = 0; x < kSsBatch; ++x) {
            const uint32_t e = e0 + 64 * x + lane_id();
            vb[x] = e < ee ? __builtin_nontemporal_load(p.col + e) : 0xFFFFFFFFu;
            nd[x] = e < ee ? du + (int32_t)__builtin_nontemporal_load(p.w + e) : 0;
          }
          sssp_relax_batch(p, vb, nd, bound, r + 2, qnear, cnear, hnear, hcnear, epoch + 1, qfar, cfar);
        }
      }
      // light near queue: 256 vertices per workgroup grab, wave64 fine-grained
      const uint32_t *qin = p.nq + (size_t)(r & 1) * p.n;
      uint32_t *wsc = p.ctr + (3 + r % 3) * kCtr;
      for (;;) {
        if (threadIdx.x == 0) s_base = add_relaxed(wsc, (uint32_t)kWg);
        __syncthreads();
        const uint32_t base = s_base;
        __syncthreads();
        if (base >= nin) break;
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < nin;
        const uint32_t u = valid ? qin[i] : 0u;
        uint32_t b = 0, deg = 0;
        int32_t du = 0;
        if (valid) { b = p.row[u]; deg = p.row[u + 1] - b; du = ld_i32(p.dist + u); }
        wave_expand<kSsBatch>(b, deg, [&](const uint32_t (&e)[kSsBatch], const bool (&ev)[kSsBatch],
                                          const uint32_t (&own)[kSsBatch]) {
          uint32_t vb[kSsBatch];
          int32_t nd[kSsBatch];
#pragma unroll
          for (int x = 0; x < kSsBatch; ++x) {
            const int32_t odu = __shfl(du, (int)own[x]);
            vb[x] = ev[x] ? __builtin_nontemporal_load(p.col + e[x]) : 0xFFFFFFFFu;
            nd[x] = ev[x] ? odu + (int32_t)__builtin_nontemporal_load(p.w + e[x]) : 0;
          }
          sssp_relax_batch(p, vb, nd, bound, r + 2, qnear, cnear, hnear, hcnear, epoch + 1, qfar, cfar);
        });
      }
      ++relax_rounds;
    } else {
      // near input empty: advance the bucket and split the far list
      // (remove_dups_dev + curdelta += DELTA, kernel_gsrb.cu:310-348,669-677)
      const uint32_t nfar = ld_relaxed(p.ctr + (6 + epoch % 3) * kCtr);
      if (nfar == 0) break;
      const uint32_t *fin = p.fq + (size_t)(epoch % 3) * p.n;
      uint32_t *fout = p.fq + (size_t)((epoch + 1) % 3) * p.n;
      uint32_t *cfout = p.ctr + (6 + (epoch + 1) % 3) * kCtr;
      if (gtid() == 0) st_relaxed(p.ctr + (6 + (epoch + 2) % 3) * kCtr, 0u);
      const int32_t nbound = (int32_t)min((int64_t)(epoch + 2) * p.delta, (int64_t)kInf - 1);
      const uint32_t span = (nfar + kWg * kSsBatch - 1) / (kWg * kSsBatch) * (kWg * kSsBatch);
      for (uint32_t i0 = blockIdx.x * kWg * kSsBatch + threadIdx.x; i0 < span; i0 += gthreads() * kSsBatch) {
        uint32_t vb[kSsBatch];
        bool to_near[kSsBatch], to_far[kSsBatch];
#pragma unroll
        for (int x = 0; x < kSsBatch; ++x) {
          const uint32_t i = i0 + x * kWg;
          vb[x] = i < nfar ? fin[i] : 0u;
          to_near[x] = to_far[x] = false;
          if (i < nfar) {
            if (ld_i32(p.dist + vb[x]) <= nbound)
              to_near[x] = __hip_atomic_exchange(p.nstamp + vb[x], r + 2, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) != r + 2;
            else
              to_far[x] = __hip_atomic_exchange(p.fstamp + vb[x], epoch + 2, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT) != epoch + 2;
          }
        }
        sssp_push_near(p, to_near, vb, qnear, cnear, hnear, hcnear);
        wave_push_batch<kSsBatch>(fout, cfout, to_far, vb);
      }
      ++epoch;
    }
    if (gb.sync<KIND>()) return;
  }
  if (gtid() == 0 && p.stats) { p.stats[0] = relax_rounds; p.stats[1] = epoch; }
}

// ===========================================================================
// PageRank, FP32, pull over the transposed CSR in a fixed summation order.
//   x'[v] = (1-a)/n + a*(D/n + sum_{u->v} x[u]/deg(u)),  D = dangling mass
// Light vertices (in-degree < kHeavy) are pulled by blocks of 64 consecutive
// vertices, one wave per block: the block's in-edges are walked as one
// flattened sequence, 64 per step (coalesced column loads), and summed per
// vertex by a segmented wave scan; the step sums are added in step order.
// Heavy vertices are cut into chunks of kChunk in-edges; a wave sums a chunk,
// and the wave that finishes the vertex's last chunk of the iteration (a
// per-vertex arrival counter) adds its chunk sums in chunk order and finalizes
// the vertex.  Every vertex is finalized in the same phase as its sums, so an
// iteration needs one grid barrier.  The dangling mass and the L1 change are
// summed in 2^-60 fixed point (integer sums: the same bits whatever the grid
// or the order).  Every order is fixed by the graph alone, so the ranks do not
// depend on the grid.
// Chunk slots: the chunk c of heavy vertex v lives at slot ceil(rrow[v]/kHeavy)+c.
// A vertex with d >= kHeavy in-edges has at most d/kHeavy chunks, so its slots
// stay below ceil(rrow[v+1]/kHeavy) and never meet another vertex's; slot i is
// owned by the vertex whose in-edge range holds edge i*kHeavy.
struct PrArgs {
  uint32_t n;
  const uint32_t *row, *rrow, *rcol;
  float alpha, tol;
  int max_iter;
  float *rank;      // output
  float *x;         // [2][n] iterates
  float *contrib;   // [2][n]
  float *hsum;      // [nslots] heavy chunk sums
  uint64_t *hent;   // [nslots] slot owner v | chunk << 32, or ~0 for an unused slot
  uint32_t nslots;
  uint32_t *harr;   // [nslots] per heavy vertex (at its first slot): chunk arrivals so far
  uint64_t *lpre;   // [ceil(n/64)] light block weight prefix inside each workgroup range
  uint64_t *spre;   // [nslots] slot weight prefix inside each workgroup range
  uint64_t *wtot;   // [2][W] range totals: light blocks, slots
  uint64_t *part;   // [2][W][2] dangling mass, L1 change (2^-60 fixed point)
  uint32_t *stats;
  uint32_t ng, nx;
};

__device__ __forceinline__ uint32_t pr_nchunks(uint32_t d) { return (d + kChunk - 1) / kChunk; }

// vertex whose in-edge range [rrow[v], rrow[v+1]) holds edge e (e < rrow[n])
__device__ uint32_t pr_edge_owner(const uint32_t *rrow, uint32_t n, uint32_t e) {
  uint32_t lo = 0, hi = n;  // rrow[lo] <= e < rrow[hi]
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (rrow[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// One wave, 64 consecutive vertices: lane j brings its vertex's first in-edge
// `eb` and run length `len` (0 for heavy, empty or out-of-range vertices) and
// returns the sum of the contributions over that run.
__device__ __forceinline__ float pr_light_block(const PrArgs &p, const float *cin, uint32_t eb,
                                                uint32_t len) {
  return wave_run_gather_sum<kPrSteps>(p.rcol, cin, eb, len);
}

__device__ __forceinline__ uint64_t q60(double v) { return (uint64_t)(v * 0x1p60); }

__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v) {
  __shared__ uint64_t red[kWg / 64];
  for (int k = 32; k >= 1; k >>= 1) v += shfl64(v, lane_id() ^ k);
  if (lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = 0;
  for (int w = 0; w < kWg / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// sum of the workgroups' fixed-point partials q of one parity, as a double
__device__ double pr_fold(const uint64_t *part, uint32_t parity, int q) {
  uint64_t acc = 0;
  for (uint32_t w = threadIdx.x; w < gridDim.x; w += kWg) acc += part[((size_t)parity * gridDim.x + w) * 2 + q];
  return (double)block_sum_u64(acc) * 0x1p-60;
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, uint32_t d) {
  const uint32_t lo = __shfl_up((uint32_t)v, d), hi = __shfl_up((uint32_t)(v >> 32), d);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint64_t up = shfl_up64(v, d);
    if (lane_id() >= d) v += up;
  }
  return v;
}

// inclusive scan over the workgroup (thread order); *total = workgroup sum
__device__ uint64_t block_incl_scan_u64(uint64_t v, uint64_t *total) {
  __shared__ uint64_t wsum[kWg / 64];
  const uint32_t wv = threadIdx.x >> 6;
  v = wave_incl_scan_u64(v);
  if (lane_id() == 63) wsum[wv] = v;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (uint32_t k = 0; k < kWg / 64; ++k) {
    if (k < wv) off += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return v + off;
}

__device__ __forceinline__ uint32_t wg_range(uint32_t count, uint32_t w) {
  return (uint32_t)((uint64_t)count * w / gridDim.x);
}

// Sum of the per-workgroup totals of a unit space (one wave).
__device__ uint64_t pr_space_total(const uint64_t *wtot) {
  uint64_t acc = 0;
  for (uint32_t j = lane_id(); j < gridDim.x; j += 64) acc += wtot[j];
  return shfl64(wave_incl_scan_u64(acc), 63);
}

// First unit whose exclusive weight prefix is >= t, in a unit space cut into
// per-workgroup ranges (pre = inclusive prefix inside each range, wtot = range
// totals).  Weights are >= 1, so the prefix is strictly increasing.  One wave.
__device__ uint32_t pr_split(const uint64_t *pre, const uint64_t *wtot, uint32_t count, uint64_t t) {
  if (t == 0) return 0;
  const uint32_t W = gridDim.x;
  uint64_t before = 0;  // total of the ranges before g
  uint32_t g = W;
  for (uint32_t base = 0; base < W; base += 64) {
    const uint32_t j = base + lane_id();
    const uint64_t v = j < W ? wtot[j] : 0;
    const uint64_t inc = wave_incl_scan_u64(v) + before;
    const uint64_t hit = ballot(j < W && inc >= t);
    if (hit) {
      const uint32_t l = (uint32_t)__builtin_ctzll(hit);
      g = base + l;
      before = shfl64(inc - v, l);
      break;
    }
    before = shfl64(inc, 63);
  }
  if (g >= W) return count;
  uint32_t lo = wg_range(count, g), hi = wg_range(count, g + 1) - 1;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (before + pre[mid] >= t) hi = mid; else lo = mid + 1;
  }
  return lo + 1;
}

template <int KIND>
__global__ void __launch_bounds__(kWg) __attribute__((amdgpu_waves_per_eu(8))) k_pagerank(ArenaView a, PrArgs p) {
  GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nx);
  const float inv_n = 1.0f / (float)p.n;
  const uint32_t m = p.rrow[p.n];
  const uint32_t nblk = (p.n + 63) / 64;
  // iteration 0 state: x = 1/n, contributions and dangling mass of x
  uint64_t dang = 0;
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) {
    p.x[v] = inv_n;
    const uint32_t d = p.row[v + 1] - p.row[v];
    p.contrib[v] = d ? inv_n / (float)d : 0.f;
    if (!d) dang += q60(inv_n);
  }
  for (uint32_t i = gtid(); i < p.nslots; i += gthreads()) p.harr[i] = 0;
  // Work units and their weights (a cost model in memory round trips): light
  // blocks of 64 vertices (1.5 x light in-edges + 256) and chunk slots (the
  // chunk's in-edges + 256, weight 1 for an unused slot).  Each workgroup prefix-sums a contiguous
  // range of each space; after the barrier every wave takes the units whose
  // exclusive prefix falls in its equal share of the total.
  __shared__ uint64_t s_w[kWg];
  {
    const uint32_t b0 = wg_range(nblk, blockIdx.x), b1 = wg_range(nblk, blockIdx.x + 1);
    uint64_t carry = 0;
    for (uint32_t cb = b0; cb < b1; cb += kWg) {
      const uint32_t nb = min((uint32_t)kWg, b1 - cb);
      for (uint32_t k = threadIdx.x >> 6; k < nb; k += kWg / 64) {
        const uint32_t v = (cb + k) * 64 + lane_id();
        uint32_t len = 0;
        if (v < p.n) {
          const uint32_t d = p.rrow[v + 1] - p.rrow[v];
          len = d < kHeavy ? d : 0u;
        }
        // cost model: a pass of 512 positions costs two memory round trips, as
        // does a 512-edge pass of a chunk (below), plus the shuffles of the
        // segmented sums (half as much again)
        const uint32_t tot = wave_sum(len);
        if (lane_id() == 0) s_w[k] = (uint64_t)tot + tot / 2 + 256;
      }
      __syncthreads();
      uint64_t tot;
      const uint64_t inc = block_incl_scan_u64(threadIdx.x < nb ? s_w[threadIdx.x] : 0, &tot);
      if (threadIdx.x < nb) p.lpre[cb + threadIdx.x] = carry + inc;
      carry += tot;
    }
    if (threadIdx.x == 0) p.wtot[blockIdx.x] = carry;
  }
  {
    // chunk slot owners: slot i belongs to chunk c of the vertex v holding edge
    // i*kHeavy when v is heavy and c = i - ceil(rrow[v]/kHeavy) < chunks(v)
    const uint32_t s0 = wg_range(p.nslots, blockIdx.x), s1 = wg_range(p.nslots, blockIdx.x + 1);
    uint64_t carry = 0;
    for (uint32_t cs = s0; cs < s1; cs += kWg) {
      const uint32_t i = cs + threadIdx.x;
      uint64_t wgt = 0;
      if (i < s1) {
        uint64_t ent = ~0ull;
        wgt = 1;
        const uint64_t e = (uint64_t)i * kHeavy;
        if (e < m) {
          const uint32_t v = pr_edge_owner(p.rrow, p.n, (uint32_t)e);
          const uint32_t rb = p.rrow[v], d = p.rrow[v + 1] - rb;
          const uint32_t first = (rb + kHeavy - 1) / kHeavy;
          if (d >= kHeavy && i >= first && i - first < pr_nchunks(d)) {
            const uint32_t c = i - first;
            ent = v | ((uint64_t)c << 32);
            wgt = (uint64_t)min(kChunk, d - c * kChunk) + 256;
          }
        }
        p.hent[i] = ent;
      }
      uint64_t tot;
      const uint64_t inc = block_incl_scan_u64(wgt, &tot);
      if (i < s1) p.spre[i] = carry + inc;
      carry += tot;
    }
    if (threadIdx.x == 0) p.wtot[gridDim.x + blockIdx.x] = carry;
  }
  dang = block_sum_u64(dang);
  if (threadIdx.x == 0) p.part[((size_t)0 * gridDim.x + blockIdx.x) * 2 + 0] = dang;
  if (gb.sync<KIND>()) return;
  // this wave's share of each unit space
  uint32_t b_lo, b_hi, s_lo, s_hi;
  {
    const uint64_t W = gwaves(), w = gwave();
    const uint64_t TL = pr_space_total(p.wtot), TS = pr_space_total(p.wtot + gridDim.x);
    b_lo = pr_split(p.lpre, p.wtot, nblk, TL * w / W);
    b_hi = pr_split(p.lpre, p.wtot, nblk, TL * (w + 1) / W);
    s_lo = pr_split(p.spre, p.wtot + gridDim.x, p.nslots, TS * w / W);
    s_hi = pr_split(p.spre, p.wtot + gridDim.x, p.nslots, TS * (w + 1) / W);
  }
  int it = 0;
  double l1 = 0;
  uint32_t cur = 0;
  for (; it < p.max_iter; ++it) {
    const float *cin = p.contrib + (size_t)cur * p.n;
    float *cout = p.contrib + (size_t)(cur ^ 1u) * p.n;
    const float *xin = p.x + (size_t)cur * p.n;
    float *xout = p.x + (size_t)(cur ^ 1u) * p.n;
    const double D = pr_fold(p.part, (uint32_t)it & 1u, 0);
    const float base = (float)((1.0 - (double)p.alpha) / p.n + (double)p.alpha * D / p.n);
    uint64_t l1q = 0, dq = 0;  // this thread's share of the L1 change and dangling mass
    // y, L1 term, next contribution and dangling term of vertex v from its sum
    auto finalize = [&](uint32_t v, float sum) {
      const float y = base + p.alpha * sum;
      l1q += q60(fabs((double)y - (double)xin[v]));
      xout[v] = y;
      const uint32_t d = p.row[v + 1] - p.row[v];
      cout[v] = d ? y / (float)d : 0.f;
      if (!d) dq += q60(y);
    };
    // heavy chunks, fixed order inside each chunk; the last chunk of a vertex
    // to finish finalizes it
    for (uint32_t sb = s_lo; sb < s_hi; sb += 64) {
      const uint32_t i = sb + lane_id();
      const uint64_t mine = i < s_hi ? p.hent[i] : ~0ull;
      uint64_t used = ballot(mine != ~0ull);
      while (used) {
        const uint32_t l = (uint32_t)__builtin_ctzll(used);
        used &= used - 1;
        const uint64_t ent = shfl64(mine, l);
        const uint32_t v = (uint32_t)ent, c = (uint32_t)(ent >> 32);
        const uint32_t rb = p.rrow[v], re = p.rrow[v + 1];
        const uint32_t eb = rb + c * kChunk, ee = min(re, eb + kChunk);
        // passes of kPrHeavySteps x 64 edges: the column loads, then the
        // gathers, in flight per lane; the lane's sum keeps edge order
        float sum = 0.f;
        for (uint32_t e = eb + lane_id(); e < ee; e += kPrHeavySteps * 64) {
          uint32_t cs[kPrHeavySteps];
#pragma unroll
          for (int u = 0; u < kPrHeavySteps; ++u) cs[u] = e + 64 * u < ee ? p.rcol[e + 64 * u] : 0xFFFFFFFFu;
          float gs[kPrHeavySteps];
#pragma unroll
          for (int u = 0; u < kPrHeavySteps; ++u) gs[u] = cs[u] != 0xFFFFFFFFu ? cin[cs[u]] : 0.f;
#pragma unroll
          for (int u = 0; u < kPrHeavySteps; ++u) sum += gs[u];
        }
        for (int k = 32; k >= 1; k >>= 1) sum += __shfl_xor(sum, k);
        if (lane_id() == 0) {
          // publish the chunk sum (agent scope: other XCDs read it), then arrive
          __hip_atomic_store(p.hsum + sb + l, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          wait_stores();
          const uint32_t first = (rb + kHeavy - 1) / kHeavy, nc = pr_nchunks(re - rb);
          const uint32_t old = add_relaxed(p.harr + first, 1u);
          if (old + 1u == (uint32_t)(it + 1) * nc) {
            float t = 0.f;
            for (uint32_t cc = 0; cc < nc && first + cc < p.nslots; ++cc)
              t += __hip_atomic_load(p.hsum + first + cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            finalize(v, t);
          }
        }
      }
    }
    // light blocks, one wave per block of 64 vertices (the next block's
    // offsets are loaded while this one is summed)
    uint32_t nb_eb = 0, nb_ee = 0;
    if (b_lo < b_hi) {
      const uint32_t v = b_lo * 64 + lane_id();
      if (v < p.n) { nb_eb = p.rrow[v]; nb_ee = p.rrow[v + 1]; }
    }
    for (uint32_t blk = b_lo; blk < b_hi; ++blk) {
      const uint32_t v = blk * 64 + lane_id();
      const bool valid = v < p.n;
      const uint32_t eb = nb_eb, din = nb_ee - nb_eb;
      if (blk + 1 < b_hi) {
        const uint32_t vn = v + 64;
        nb_eb = vn < p.n ? p.rrow[vn] : 0u;
        nb_ee = vn < p.n ? p.rrow[vn + 1] : 0u;
      }
      const bool light = din < kHeavy;
      const float sum = pr_light_block(p, cin, eb, light ? din : 0u);
      if (valid && light) finalize(v, sum);
    }
    l1q = block_sum_u64(l1q);
    dq = block_sum_u64(dq);
    if (threadIdx.x == 0) {
      p.part[((size_t)(((uint32_t)it + 1) & 1u) * gridDim.x + blockIdx.x) * 2 + 0] = dq;
      p.part[((size_t)((uint32_t)it & 1u) * gridDim.x + blockIdx.x) * 2 + 1] = l1q;
    }
    if (gb.sync<KIND>()) return;
    cur ^= 1u;
    l1 = pr_fold(p.part, (uint32_t)it & 1u, 1);
    if (l1 < (double)p.tol) { ++it; break; }
  }
  const float *xf = p.x + (size_t)cur * p.n;
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) p.rank[v] = xf[v];
  if (gtid() == 0 && p.stats) {
    p.stats[0] = (uint32_t)it;
    const float l1f = (float)l1;
    p.stats[1] = __float_as_uint(l1f);
  }
}

#define HS_KIND_SWITCH(KERNEL, kind)                                            \
  [&]() -> const void * {                                                       \
    switch (kind) {                                                             \
      case HS_BAR_FLAT: return (const void *)KERNEL<HS_BAR_FLAT>;               \
      case HS_BAR_TWO_LEVEL: return (const void *)KERNEL<HS_BAR_TWO_LEVEL>;     \
      case HS_BAR_XCD: return (const void *)KERNEL<HS_BAR_XCD>;                 \
      case HS_BAR_XCD_FLAT_RELEASE: return (const void *)KERNEL<HS_BAR_XCD_FLAT_RELEASE>; \
      case HS_BAR_TREE: return (const void *)KERNEL<HS_BAR_TREE>;               \
      case HS_BAR_LF: return (const void *)KERNEL<HS_BAR_LF>;                   \
      case HS_BAR_COOP: return (const void *)KERNEL<HS_BAR_COOP>;               \
      case HS_BAR_NAIVE_ALL: return (const void *)KERNEL<HS_BAR_NAIVE_ALL>;     \
      default: return nullptr;                                                  \
    }                                                                           \
  }()

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int grid_for(const void *k, hs_arena *ar, int num_wgs, int per_cu = 2) {
  if (num_wgs <= 0) {  // the default grid, within the arena's barrier slots
    num_wgs = hs_host::default_grid(k, kWg, per_cu);
    if (num_wgs > ar->view.max_blocks) num_wgs = ar->view.max_blocks;
  }
  if (num_wgs > ar->view.max_blocks) return -1;
  if (hs_host::check_residency(k, num_wgs, kWg, 0) != HS_OK) return -1;
  return num_wgs;
}

uint64_t ms_hcap(uint32_t n, uint64_t m) { return m / kHeavy + m / kMsChunk + 64; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------- BFS
// heavy chunk list capacity: a vertex of d >= kHeavy out-edges has at most
// d / kHeavy chunks of kBfsChunk, and each vertex is reached once
static uint64_t bfs_hcap(uint64_t m) { return m / kHeavy + 64; }

int hs_bfs_work_size(uint32_t n, uint64_t m, uint64_t *bytes) {
  if (!bytes) return HS_ERR_INVALID;
  *bytes = align256(sizeof(uint32_t) * 2ull * n) + align256(sizeof(uint32_t) * 12 * kCtr) +
           2 * align256(sizeof(uint64_t) * 2 * bfs_hcap(m));
  return HS_OK;
}

static int bfs_launch(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
                      int32_t src, int32_t *d_level, void *d_work, int kind, int num_wgs, bool lb,
                      int degree_limit, uint32_t *d_stats, void *stream) {
  if (!ar || !d_row || (m && !d_col) || !d_level || !d_work || n == 0 || src < 0 ||
      (uint32_t)src >= n || m > 0xFFFFFFFFull || (lb && degree_limit != 0 && degree_limit < (int)kHeavy))
    return HS_ERR_INVALID;
  const void *k = lb ? HS_KIND_SWITCH(k_bfs_lb, kind) : HS_KIND_SWITCH(k_bfs, kind);
  if (!k) return HS_ERR_INVALID;
  const int g = grid_for(k, ar, num_wgs);
  if (g < 0) return HS_ERR_RESIDENCY;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::prepare_barrier(ar, kind, s);
  if (rc) return rc;
  BfsArgs p{};
  p.n = n; p.row = d_row; p.col = d_col; p.src = src; p.level = d_level;
  p.q = (uint32_t *)d_work;
  p.ctr = (uint32_t *)((char *)d_work + align256(sizeof(uint32_t) * 2ull * n));
  p.hq = (uint64_t *)((char *)p.ctr + align256(sizeof(uint32_t) * 12 * kCtr));
  p.hcap = bfs_hcap(m);
  p.lq = (uint64_t *)((char *)p.hq + align256(sizeof(uint64_t) * 2 * p.hcap));
  p.stats = d_stats;
  // the reference's DEGREE_LIMIT under -l is the grid's thread count
  // (LONESTAR/rt/include/ggcuda.h:19-22); at least kHeavy, so that the heavy
  // list of a level (at most m / limit entries) fits hcap
  p.lb_limit = lb ? (degree_limit > 0 ? (uint32_t)degree_limit : (uint32_t)g * kWg) : 0u;
  p.ng = (uint32_t)ar->view.num_groups;
  p.nx = (uint32_t)hs_host::xcd_groups(g);
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_host::launch_persistent(k, args, kind, g, kWg, 0, s);
}

int hs_bfs(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
           int32_t src, int32_t *d_level, void *d_work, int kind, int num_wgs, uint32_t *d_stats,
           void *stream) {
  return bfs_launch(ar, n, m, d_row, d_col, src, d_level, d_work, kind, num_wgs, false, 0, d_stats, stream);
}

int hs_bfs_lb(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
              int32_t src, int32_t *d_level, void *d_work, int kind, int num_wgs, int degree_limit,
              uint32_t *d_stats, void *stream) {
  return bfs_launch(ar, n, m, d_row, d_col, src, d_level, d_work, kind, num_wgs, true, degree_limit, d_stats,
                    stream);
}

// ---------------------------------------------------------------- MS-BFS
int hs_msbfs_work_size(uint32_t n, uint64_t m, int nsrc, uint64_t *bytes) {
  (void)nsrc;
  if (!bytes) return HS_ERR_INVALID;
  const uint64_t hcap = ms_hcap(n, m);
  *bytes = 3 * align256(sizeof(uint64_t) * n) + align256(sizeof(uint64_t) * 4 * hcap) +
           align256(sizeof(uint32_t) * kMsCounters * kCtr) + align256(sizeof(uint64_t) * 2 * 4 * 4096) +
           align256(sizeof(uint64_t) * kMsPlanes * (uint64_t)n) + 2 * align256(sizeof(uint64_t) * ((n + 63) / 64)) +
           align256(HS_MSBFS_MAX_LEVELS);
  return HS_OK;
}

int hs_msbfs(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
             const uint32_t *d_rrow, const uint32_t *d_rcol, const int32_t *d_srcs, int nsrc,
             int32_t *d_levels, void *d_work, int kind, int num_wgs, int schedule, const float *d_mlp,
             const uint8_t *h_forced, uint32_t *d_trace, void *stream) {
  if (!ar || !d_row || !d_rrow || (m && (!d_col || !d_rcol)) || !d_srcs || nsrc <= 0 || nsrc > 64 ||
      !d_levels || !d_work || n == 0 || m > 0xFFFFFFFFull || schedule < 0 || schedule > 4 ||
      (schedule == 3 && !d_mlp) || (schedule == 4 && !h_forced))
    return HS_ERR_INVALID;
  const void *k = HS_KIND_SWITCH(k_msbfs, kind);
  if (!k) return HS_ERR_INVALID;
  // 4 WGs/CU (the kernel's occupancy): at R-MAT scale 22 with 64 sources,
  // 1.95 ms against 2.41 ms at 2 WGs/CU (the pull's gathers and the update
  // passes both gain from the waves in flight)
  const int g = grid_for(k, ar, num_wgs, 4);
  if (g < 0 || g > 4096) return HS_ERR_RESIDENCY;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::prepare_barrier(ar, kind, s);
  if (rc) return rc;
  MsArgs p{};
  p.n = n; p.m = m; p.row = d_row; p.col = d_col; p.rrow = d_rrow; p.rcol = d_rcol;
  p.srcs = d_srcs; p.nsrc = nsrc; p.levels = d_levels;
  char *w = (char *)d_work;
  p.fb[0] = (uint64_t *)w; w += align256(sizeof(uint64_t) * n);
  p.fb[1] = (uint64_t *)w; w += align256(sizeof(uint64_t) * n);
  p.vis = (uint64_t *)w; w += align256(sizeof(uint64_t) * n);
  p.hcap = ms_hcap(n, m);
  p.hlist = (uint64_t *)w; w += align256(sizeof(uint64_t) * 4 * p.hcap);
  p.hcnt = (uint32_t *)w; w += align256(sizeof(uint32_t) * kMsCounters * kCtr);
  p.part = (uint64_t *)w; w += align256(sizeof(uint64_t) * 2 * 4 * 4096);
  p.planes = (uint64_t *)w; w += align256(sizeof(uint64_t) * kMsPlanes * (uint64_t)n);
  p.fbm[0] = (uint64_t *)w; w += align256(sizeof(uint64_t) * ((n + 63) / 64));
  p.fbm[1] = (uint64_t *)w; w += align256(sizeof(uint64_t) * ((n + 63) / 64));
  uint8_t *forced = (uint8_t *)w;
  if (schedule == 4) HS_TRY_HIP(hipMemcpyAsync(forced, h_forced, HS_MSBFS_MAX_LEVELS, hipMemcpyHostToDevice, s));
  p.schedule = schedule; p.mlp = d_mlp; p.forced = forced; p.trace = d_trace;
  p.ng = (uint32_t)ar->view.num_groups;
  p.nxcd = (uint32_t)hs_host::xcd_groups(g);
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_host::launch_persistent(k, args, kind, g, kWg, 0, s);
}

// ---------------------------------------------------------------- SSSP
int hs_sssp_work_size(uint32_t n, uint64_t m, uint64_t *bytes) {
  if (!bytes) return HS_ERR_INVALID;
  *bytes = align256(sizeof(uint32_t) * 2ull * n) + align256(sizeof(uint32_t) * 3ull * n) +
           2 * align256(sizeof(uint32_t) * (uint64_t)n) + align256(sizeof(uint32_t) * 12 * kCtr) +
           align256(sizeof(uint64_t) * 2 * bfs_hcap(m));
  return HS_OK;
}

int hs_sssp(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
            const uint32_t *d_w, int32_t src, int32_t delta, int32_t *d_dist, void *d_work, int kind,
            int num_wgs, uint32_t *d_stats, void *stream) {
  if (!ar || !d_row || (m && (!d_col || !d_w)) || !d_dist || !d_work || n == 0 || src < 0 ||
      (uint32_t)src >= n || m > 0xFFFFFFFFull)
    return HS_ERR_INVALID;
  const void *k = HS_KIND_SWITCH(k_sssp, kind);
  if (!k) return HS_ERR_INVALID;
  const int g = grid_for(k, ar, num_wgs);
  if (g < 0) return HS_ERR_RESIDENCY;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::prepare_barrier(ar, kind, s);
  if (rc) return rc;
  SsspArgs p{};
  p.n = n; p.row = d_row; p.col = d_col; p.w = d_w; p.src = src;
  p.delta = delta > 0 ? delta : 10000;  // DELTA default, LONESTAR/apps/sssp/support.cu:10
  p.dist = d_dist;
  char *w = (char *)d_work;
  p.nq = (uint32_t *)w; w += align256(sizeof(uint32_t) * 2ull * n);
  p.fq = (uint32_t *)w; w += align256(sizeof(uint32_t) * 3ull * n);
  p.nstamp = (uint32_t *)w; w += align256(sizeof(uint32_t) * (uint64_t)n);
  p.fstamp = (uint32_t *)w; w += align256(sizeof(uint32_t) * (uint64_t)n);
  p.ctr = (uint32_t *)w; w += align256(sizeof(uint32_t) * 12 * kCtr);
  p.hq = (uint64_t *)w;
  p.hcap = bfs_hcap(m);
  p.stats = d_stats;
  p.ng = (uint32_t)ar->view.num_groups;
  p.nx = (uint32_t)hs_host::xcd_groups(g);
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_host::launch_persistent(k, args, kind, g, kWg, 0, s);
}

// ---------------------------------------------------------------- PageRank
// heavy chunk slots: ceil(m / kHeavy) (see PrArgs)
static uint64_t pr_slots(uint64_t m) { return (m + kHeavy - 1) / kHeavy; }

int hs_pagerank_work_size(uint32_t n, uint64_t m, uint64_t *bytes) {
  if (!bytes) return HS_ERR_INVALID;
  const uint64_t hc = pr_slots(m) + 1, nblk = (n + 63ull) / 64;
  *bytes = align256(sizeof(float) * 2ull * n) * 2 + align256(sizeof(uint64_t) * 2 * 2 * 4096) +
           align256(sizeof(uint64_t) * hc) * 2 + align256(sizeof(float) * hc) +
           align256(sizeof(uint32_t) * hc) + align256(sizeof(uint64_t) * nblk) +
           align256(sizeof(uint64_t) * 2 * 4096);
  return HS_OK;
}

int hs_pagerank_f32(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row,
                    const uint32_t *d_rrow, const uint32_t *d_rcol, float alpha, float tol,
                    int max_iter, float *d_rank, void *d_work, int kind, int num_wgs,
                    uint32_t *d_stats, void *stream) {
  if (!ar || !d_row || !d_rrow || (m && !d_rcol) || !d_rank || !d_work || n == 0 || !(alpha > 0.f) ||
      !(alpha < 1.f) || max_iter < 1 || m > 0xFFFFFFFFull)
    return HS_ERR_INVALID;
  const void *k = HS_KIND_SWITCH(k_pagerank, kind);
  if (!k) return HS_ERR_INVALID;
  const int g = grid_for(k, ar, num_wgs, 4);  // 4 WGs/CU: latency-bound gathers
  if (g < 0 || g > 4096) return HS_ERR_RESIDENCY;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::prepare_barrier(ar, kind, s);
  if (rc) return rc;
  PrArgs p{};
  p.n = n; p.row = d_row; p.rrow = d_rrow; p.rcol = d_rcol; p.alpha = alpha; p.tol = tol;
  p.max_iter = max_iter; p.rank = d_rank;
  const uint64_t hc = pr_slots(m) + 1;
  char *w = (char *)d_work;
  p.x = (float *)w; w += align256(sizeof(float) * 2ull * n);
  p.contrib = (float *)w; w += align256(sizeof(float) * 2ull * n);
  p.part = (uint64_t *)w; w += align256(sizeof(uint64_t) * 2 * 2 * 4096);
  p.hent = (uint64_t *)w; w += align256(sizeof(uint64_t) * hc);
  p.spre = (uint64_t *)w; w += align256(sizeof(uint64_t) * hc);
  p.hsum = (float *)w; w += align256(sizeof(float) * hc);
  p.harr = (uint32_t *)w; w += align256(sizeof(uint32_t) * hc);
  p.lpre = (uint64_t *)w; w += align256(sizeof(uint64_t) * ((n + 63ull) / 64));
  p.wtot = (uint64_t *)w;
  p.nslots = (uint32_t)pr_slots(m);
  p.stats = d_stats;
  p.ng = (uint32_t)ar->view.num_groups;
  p.nx = (uint32_t)hs_host::xcd_groups(g);
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_host::launch_persistent(k, args, kind, g, kWg, 0, s);
}

}  // extern "C"
