// mssp.hip -- up to 64 single-source shortest-path problems in one persistent
// launch (the batched form of LONESTAR/apps/sssp/kernel_gsrb.cu, whose
// near/far relaxation with atomicMin runs one source per launch).
//
// Layout: working distances are vertex-major, dist[v][s] for s < 64, so the 64
// distances of one vertex are 256 contiguous bytes and one wave64 relaxes an
// edge for all sources at once (lane s = source s): one 256-byte load of the
// head's distances, a compare, and atomicMin only in the lanes that improve.
// Per vertex, 64-bit masks say which sources changed it: `nx` collects the
// improvements of the current round, `pend` the changes not yet expanded.
//
// Rounds (grid barrier between phases, any hs_barrier_kind):
//   A  every vertex folds nx into pend; vertices whose pending sources have a
//      distance below the round's bound B are expanded this round (cur), the
//      rest stay pending; expanded vertices go to a work list (vertices of
//      degree >= kHeavy as chunks of kChunk edges); totals of expanded and
//      still-pending vertices and the smallest pending distance per workgroup.
//   B  waves take list entries (work stealing) and relax their edges for the
//      sources in cur[v].
// B starts at delta and grows to (smallest pending distance + delta) whenever
// nothing is below it: delta-stepping (Meyer & Sanders 2003) with one bound
// shared by all sources; delta = INT32_MAX is plain frontier Bellman-Ford.  The
// relaxation is monotone and every improvement is re-expanded, so the result is
// the exact shortest-path distance whatever the order: bit-exact against
// Dijkstra.  A final pass transposes dist into the caller's [nsrc][n] array.
#include <hip/hip_runtime.h>

#include "hs_graph.h"
#include "hs_internal.hpp"
#include "hs_wave.hpp"

using namespace hs;

namespace hs_host {
int prepare_barrier(hs_arena *ar, int kind, hipStream_t s);
int launch_persistent(const void *k, void **args, int kind, int num_wgs, int wg_size, size_t lds,
                      hipStream_t s);
}  // namespace hs_host

namespace {

constexpr int kWg = 256;
constexpr int32_t kInf = 0x7FFFFFFF;
constexpr uint32_t kHeavy = 256;   // degree at which a vertex is cut into chunks
constexpr uint32_t kChunk = 1024;  // edges per chunk of a heavy vertex
constexpr int kCtr = 64;           // words between counters
constexpr int kEdgeBatch = 8;      // edges whose head distances are in flight per wave
constexpr int kGrab = 8;           // list entries a wave takes per atomic

struct MsspArgs {
  uint32_t n;
  uint64_t m;
  const uint32_t *row, *col, *w;
  const int32_t *srcs;
  int nsrc;
  int32_t delta;
  int32_t *dist;        // [n][64]
  int32_t *out;         // [nsrc][n]
  uint64_t *pend, *nx, *cur;  // [n]
  uint64_t *items;      // [2][icap]: v | chunk << 32
  uint64_t icap;
  uint32_t *ctr;        // [2 parities] x {items, steal}, kCtr words apart
  uint64_t *part;       // [W][4]: expanded vertices, pending vertices, min pending distance, expanded pairs
  uint32_t *stats;      // {rounds, bound steps, expanded pairs (low 32 bits), 0}
  uint32_t ng, nxcd;
};

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * kWg + threadIdx.x; }
__device__ __forceinline__ uint32_t gthreads() { return gridDim.x * kWg; }
__device__ __forceinline__ uint32_t *items_cnt(const MsspArgs &p, uint32_t par) { return p.ctr + (2 * par) * kCtr; }
__device__ __forceinline__ uint32_t *steal_cnt(const MsspArgs &p, uint32_t par) { return p.ctr + (2 * par + 1) * kCtr; }

__device__ __forceinline__ int32_t ld_i32(const int32_t *q) {
  return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Phase A over all vertices: pend |= nx; expand the pending sources below
// `bound`.  Per-workgroup totals go to part[wg].
__device__ void mssp_select(const MsspArgs &p, uint32_t par, int32_t bound) {
  uint64_t expanded = 0, pending = 0, pairs = 0;
  int32_t minpend = kInf;
  uint64_t *list = p.items + (size_t)par * p.icap;
  uint32_t *cnt = items_cnt(p, par);
  // whole waves step together (v = wave base + lane), so the list append can
  // be one atomic per wave
  for (uint32_t wb = blockIdx.x * kWg + (threadIdx.x & ~63u); wb < p.n; wb += gthreads()) {
    const uint32_t v = wb + lane_id();
    const bool valid = v < p.n;
    uint64_t pm = valid ? p.pend[v] | p.nx[v] : 0ull;
    uint64_t near = 0;
    if (pm) {
      p.nx[v] = 0;
      if (bound == kInf) {
        near = pm;
      } else {
        uint64_t bits = pm;
        while (bits) {
          const int s = __builtin_ctzll(bits);
          bits &= bits - 1;
          const int32_t d = ld_i32(p.dist + (size_t)v * 64 + s);
          if (d < bound) near |= 1ull << s;
          else minpend = d < minpend ? d : minpend;
        }
      }
      p.pend[v] = pm & ~near;
      pending += (pm & ~near) != 0;
    }
    if (valid) p.cur[v] = near;
    uint32_t nc = 0;
    if (near) {
      ++expanded;
      pairs += (uint64_t)__builtin_popcountll(near);
      const uint32_t b = p.row[v], deg = p.row[v + 1] - b;
      nc = deg >= kHeavy ? (deg + kChunk - 1) / kChunk : (deg ? 1u : 0u);
    }
    // wave-aggregated append of the list entries (one atomic per wave)
    const uint32_t incl = wave_incl_scan(nc);
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    if (total) {
      uint32_t at = 0;
      if (lane_id() == 0) at = add_relaxed(cnt, total);
      at = __shfl(at, 0) + incl - nc;
      for (uint32_t c = 0; c < nc && at + c < p.icap; ++c) list[at + c] = v | ((uint64_t)c << 32);
    }
  }
  // workgroup totals
  __shared__ uint64_t red[4][kWg / 64];
  uint64_t vals[4] = {expanded, pending, (uint64_t)(uint32_t)minpend, pairs};
  for (int q = 0; q < 4; ++q) {
    uint64_t x = vals[q];
    for (int k = 32; k >= 1; k >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)x, k), hi = __shfl_xor((uint32_t)(x >> 32), k);
      const uint64_t y = ((uint64_t)hi << 32) | lo;
      x = q == 2 ? (y < x ? y : x) : x + y;
    }
    if (lane_id() == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    uint64_t t = threadIdx.x == 2 ? (uint64_t)kInf : 0;
    for (int w = 0; w < kWg / 64; ++w) {
      const uint64_t y = red[threadIdx.x][w];
      t = threadIdx.x == 2 ? (y < t ? y : t) : t + y;
    }
    st_relaxed(reinterpret_cast<uint32_t *>(p.part + (size_t)blockIdx.x * 4 + threadIdx.x), (uint32_t)t);
    st_relaxed(reinterpret_cast<uint32_t *>(p.part + (size_t)blockIdx.x * 4 + threadIdx.x) + 1,
               (uint32_t)(t >> 32));
  }
  __syncthreads();
}

struct Totals {
  uint64_t expanded, pending, pairs;
  int32_t minpend;
};

__device__ Totals mssp_totals(const MsspArgs &p) {
  __shared__ uint64_t red[4][kWg / 64];
  uint64_t acc[4] = {0, 0, (uint64_t)kInf, 0};
  for (uint32_t w = threadIdx.x; w < gridDim.x; w += kWg)
    for (int q = 0; q < 4; ++q) {
      const uint32_t *wd = reinterpret_cast<const uint32_t *>(p.part + (size_t)w * 4 + q);
      const uint64_t y = ((uint64_t)ld_relaxed(wd + 1) << 32) | ld_relaxed(wd);
      acc[q] = q == 2 ? (y < acc[q] ? y : acc[q]) : acc[q] + y;
    }
  for (int q = 0; q < 4; ++q) {
    uint64_t x = acc[q];
    for (int k = 32; k >= 1; k >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)x, k), hi = __shfl_xor((uint32_t)(x >> 32), k);
      const uint64_t y = ((uint64_t)hi << 32) | lo;
      x = q == 2 ? (y < x ? y : x) : x + y;
    }
    if (lane_id() == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  uint64_t t[4] = {0, 0, (uint64_t)kInf, 0};
  for (int q = 0; q < 4; ++q)
    for (int w = 0; w < kWg / 64; ++w) t[q] = q == 2 ? (red[q][w] < t[q] ? red[q][w] : t[q]) : t[q] + red[q][w];
  __syncthreads();
  return Totals{t[0], t[1], t[3], (int32_t)t[2]};
}

// Phase B: waves take list entries; lane s relaxes source s.
__device__ void mssp_relax(const MsspArgs &p, uint32_t par) {
  const uint32_t lane = lane_id();
  const uint64_t *list = p.items + (size_t)par * p.icap;
  const uint32_t total = ld_relaxed(items_cnt(p, par));
  const uint32_t nitems = total < p.icap ? total : (uint32_t)p.icap;
  uint32_t *steal = steal_cnt(p, par);
  // kGrab entries per atomic: one shared counter serializes its atomics, so
  // one per entry would cost ~11 ns x entries
  for (;;) {
    uint32_t it0 = 0;
    if (lane == 0) it0 = add_relaxed(steal, (uint32_t)kGrab);
    it0 = __shfl(it0, 0);
    if (it0 >= nitems) break;
    const uint32_t itn = it0 + (uint32_t)kGrab < nitems ? it0 + (uint32_t)kGrab : nitems;
    // the grabbed entries in one load (lane k holds entry it0 + k)
    const uint64_t ents = lane < itn - it0 ? list[it0 + lane] : 0ull;
    // the next entry's mask, head distances and offsets are loaded while the
    // current one relaxes (they do not depend on each other)
    uint32_t nv = __shfl((uint32_t)ents, 0);
    uint64_t nmask = p.cur[nv];
    int32_t ndv = ld_i32(p.dist + (size_t)nv * 64 + lane);
    uint32_t nrb = p.row[nv], nre = p.row[nv + 1];
    for (uint32_t it = it0; it < itn; ++it) {
      const uint32_t v = nv;
      const uint32_t c = __shfl((uint32_t)(ents >> 32), (int)(it - it0));
      const uint64_t mask = nmask;
      const bool mine = (mask >> lane) & 1u;
      const int32_t dv = mine ? ndv : kInf;
      const uint32_t rb = nrb, re = nre;
      if (it + 1 < itn) {
        nv = __shfl((uint32_t)ents, (int)(it + 1 - it0));
        nmask = p.cur[nv];
        ndv = ld_i32(p.dist + (size_t)nv * 64 + lane);
        nrb = p.row[nv];
        nre = p.row[nv + 1];
      }
      (void)v;
      const uint32_t eb = rb + c * kChunk;
      const uint32_t ee = re - eb > kChunk ? eb + kChunk : re;
      for (uint32_t e0 = eb; e0 < ee; e0 += kEdgeBatch) {
        // lane k < kEdgeBatch fetches edge e0 + k; the batch is broadcast
        const uint32_t ek = e0 + lane;
        const bool ev = lane < (uint32_t)kEdgeBatch && ek < ee;
        const uint32_t uk = ev ? p.col[ek] : 0u, wk = ev ? p.w[ek] : 0u;
        int32_t cu[kEdgeBatch];
        uint32_t us[kEdgeBatch], ws[kEdgeBatch];
  #pragma unroll
        for (int k = 0; k < kEdgeBatch; ++k) {
          us[k] = __shfl(uk, k);
          ws[k] = __shfl(wk, k);
          cu[k] = (e0 + k < ee && mine) ? ld_i32(p.dist + (size_t)us[k] * 64 + lane) : kInf;
        }
        // the batch's atomicMins are all issued before any result is used
        int32_t old[kEdgeBatch], nd[kEdgeBatch];
        bool tried[kEdgeBatch];
  #pragma unroll
        for (int k = 0; k < kEdgeBatch; ++k) {
          tried[k] = false;
          nd[k] = 0;
          old[k] = kInf;
          if (e0 + k < ee && mine) {
            nd[k] = dv + (int32_t)ws[k];
            if (nd[k] < cu[k]) {
              tried[k] = true;
              old[k] = __hip_atomic_fetch_min(p.dist + (size_t)us[k] * 64 + lane, nd[k], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
  #pragma unroll
        for (int k = 0; k < kEdgeBatch; ++k) {
          const uint64_t im = __builtin_amdgcn_ballot_w64(tried[k] && nd[k] < old[k]);
          if (im && lane == 0)
            __hip_atomic_fetch_or(p.nx + us[k], im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  }
}

template <int KIND>
__global__ void __launch_bounds__(kWg) k_mssp(ArenaView a, MsspArgs p) {
  GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nxcd);
  // init: every distance INF, nothing pending; counters of both parities 0
  const uint64_t nd = (uint64_t)p.n * 64;
  for (uint64_t i = gtid(); i < nd; i += gthreads()) p.dist[i] = kInf;
  for (uint32_t v = gtid(); v < p.n; v += gthreads()) { p.pend[v] = 0; p.nx[v] = 0; p.cur[v] = 0; }
  if (gtid() < 4) p.ctr[gtid() * kCtr] = 0;
  if (gb.sync<KIND>()) return;
  if (blockIdx.x == 0 && threadIdx.x < (uint32_t)p.nsrc) {
    const uint32_t s = threadIdx.x, v = (uint32_t)p.srcs[s];
    p.dist[(size_t)v * 64 + s] = 0;
    __hip_atomic_fetch_or(p.nx + v, 1ull << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (gb.sync<KIND>()) return;
  int32_t bound = p.delta;
  uint32_t rounds = 0, steps = 0;
  uint64_t pairs = 0;
  for (uint32_t r = 0;; ++r) {
    const uint32_t par = r & 1u;
    mssp_select(p, par, bound);
    // the other parity's counters were last used in round r-1: clear them for r+1
    if (gtid() < 2) st_relaxed(p.ctr + (2 * (par ^ 1u) + gtid()) * kCtr, 0u);
    if (gb.sync<KIND>()) return;
    const Totals t = mssp_totals(p);
    pairs += t.pairs;
    if (t.expanded == 0) {
      if (t.pending == 0) break;
      // nothing below the bound: move it past the smallest pending distance
      bound = t.minpend > kInf - p.delta ? kInf : t.minpend + p.delta;
      ++steps;
      if (gb.sync<KIND>()) return;  // every workgroup has read the totals
      continue;
    }
    ++rounds;
    mssp_relax(p, par);
    if (gb.sync<KIND>()) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && p.stats) {
    p.stats[0] = rounds;
    p.stats[1] = steps;
    p.stats[2] = (uint32_t)pairs;
  }
  // transpose dist[v][s] -> out[s][v] through LDS, 64 vertices per tile
  __shared__ int32_t tile[64][65];
  for (uint32_t v0 = blockIdx.x * 64; v0 < p.n; v0 += gridDim.x * 64) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t idx = i * kWg + threadIdx.x, vl = idx / 64, s = idx % 64;
      tile[vl][s] = v0 + vl < p.n ? p.dist[(size_t)(v0 + vl) * 64 + s] : kInf;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t vl = threadIdx.x % 64, s = threadIdx.x / 64 + 4 * i;
      if (s < (uint32_t)p.nsrc && v0 + vl < p.n) p.out[(size_t)s * p.n + v0 + vl] = tile[vl][s];
    }
    __syncthreads();
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

uint64_t mssp_icap(uint32_t n, uint64_t m) { return (uint64_t)n + m / kChunk + 64; }

struct MsspLayout {
  size_t dist, pend, nx, cur, items, ctr, part, total;
};

MsspLayout mssp_layout(uint32_t n, uint64_t m, int max_wgs) {
  MsspLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  L.dist = take(sizeof(int32_t) * 64 * (size_t)n);
  L.pend = take(sizeof(uint64_t) * (size_t)n);
  L.nx = take(sizeof(uint64_t) * (size_t)n);
  L.cur = take(sizeof(uint64_t) * (size_t)n);
  L.items = take(sizeof(uint64_t) * 2 * mssp_icap(n, m));
  L.ctr = take(sizeof(uint32_t) * 4 * kCtr);
  L.part = take(sizeof(uint64_t) * 4 * (size_t)max_wgs);
  L.total = off;
  return L;
}

}  // namespace

extern "C" {

int hs_mssp_work_size(uint32_t n, uint64_t m, int nsrc, uint64_t *bytes) {
  if (!bytes || nsrc <= 0 || nsrc > 64) return HS_ERR_INVALID;
  *bytes = mssp_layout(n, m, 8192).total;
  return HS_OK;
}

int hs_mssp(hs_arena *ar, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
            const uint32_t *d_w, const int32_t *d_srcs, int nsrc, int32_t delta, int32_t *d_dist, void *d_work,
            int kind, int num_wgs, uint32_t *d_stats, void *stream) {
  if (!ar || !d_row || !d_col || !d_w || !d_srcs || !d_dist || !d_work || n == 0 || nsrc <= 0 || nsrc > 64 ||
      delta <= 0)
    return HS_ERR_INVALID;
  const void *k = HS_KIND_SWITCH(k_mssp, kind);
  if (!k) return HS_ERR_INVALID;
  // 4 workgroups per CU (measured: 36 / 41 ms at 4 / 2 per CU for 64 sources on
  // R-MAT 20; the relaxation is latency-bound per wave)
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 4);
  if (num_wgs > ar->view.max_blocks || num_wgs > 8192) return HS_ERR_INVALID;
  if (hs_host::check_residency(k, num_wgs, kWg, 0) != HS_OK) return HS_ERR_RESIDENCY;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::prepare_barrier(ar, kind, s);
  if (rc) return rc;
  const MsspLayout L = mssp_layout(n, m, 8192);
  char *w = (char *)d_work;
  MsspArgs p{};
  p.n = n;
  p.m = m;
  p.row = d_row;
  p.col = d_col;
  p.w = d_w;
  p.srcs = d_srcs;
  p.nsrc = nsrc;
  p.delta = delta;
  p.dist = (int32_t *)(w + L.dist);
  p.out = d_dist;
  p.pend = (uint64_t *)(w + L.pend);
  p.nx = (uint64_t *)(w + L.nx);
  p.cur = (uint64_t *)(w + L.cur);
  p.items = (uint64_t *)(w + L.items);
  p.icap = mssp_icap(n, m);
  p.ctr = (uint32_t *)(w + L.ctr);
  p.part = (uint64_t *)(w + L.part);
  p.stats = d_stats;
  p.ng = (uint32_t)ar->view.num_groups;
  p.nxcd = (uint32_t)hs_host::xcd_groups(num_wgs);
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_host::launch_persistent(k, args, kind, num_wgs, kWg, 0, s);
}

}  // extern "C"
