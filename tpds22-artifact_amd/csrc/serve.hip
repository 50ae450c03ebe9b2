// serve.hip -- a serving loop that runs on the device as one persistent kernel.
//
// No reference counterpart.  The host publishes requests into a ring of
// descriptors in coherent pinned host memory; every workgroup of the
// persistent grid (1) takes a token of an admitted request from a shared queue
// and works on that request's tiles, or (2) claims the next published request
// (the 8 poller workgroups), admits it when its working-memory cost fits the
// budget -- a counting semaphore whose releases (request completions) raise
// its priority word so that waiting admissions back off, the priority rule of
// Heterosync/cudaLocksSemaphoreSpin.cu:411-622 -- and appends tokens that
// bring workgroups to it, or (3) finding neither, joins a running request
// that still has tiles through the active-request set.  Tiles of a phase are
// handed out by a fetch-and-add on the request's claim word.  Requests that
// arrive while others run are admitted and overlap on the device (continuous
// batching).  The last tile of a request writes its completion record to
// host memory.
//   reduce_i32 : tiles of 64Ki elements, sums combined with atomics (wrap)
//   scan_u32   : tiles of 16384 elements, carries by a wave-parallel decoupled
//                look-back
//   bfs        : phases: level initialization (4096 vertices per tile), then
//                one phase per BFS level, about 320 tiles of at least 256
//                frontier vertices
//   pagerank   : phases: contributions of x0 (1024 vertices per tile; these
//                tiles also place the starts of the later tiles), then one
//                phase per iteration over about 320 ranges of equal weight
//                (in-edges + 4 per vertex; a range may start or end inside a
//                vertex's in-edges) -- pull of the previous contributions,
//                then the tile's next contributions, dangling mass and L1
//                change; optionally until the L1 change is below a tolerance
// A multi-phase request runs phase by phase: the last tile of a phase to
// finish (an arrival counter: the fan-in half of a barrier, no waiting) reads
// what the phase produced, and opens the next phase or completes the
// request.  No tile ever waits for another, so requests of any size
// share the grid without deadlock.
// The kernel runs on its own non-blocking stream, every wait is bounded, and an
// abort word in host memory stops the loop.
#include <hip/hip_runtime.h>

#include <string.h>
#include <unistd.h>

#include <vector>

#include "hs_internal.hpp"
#include "hs_wave.hpp"
#include "serve_tiles.hpp"

using namespace hs;

namespace {

constexpr int kWg = 256;
constexpr int kScanSegs = 16;                   // scan: 256-element segments per wave and tile
constexpr uint32_t kScanTile = 4 * kScanSegs * 256;  // 16384 elements
constexpr uint32_t kRedTile = 65536;
constexpr uint32_t kMaxTilesPerReq = 1u << 16;  // scan: 2^30 elements per request at most
constexpr uint32_t kPrInitTile = 1024;          // pagerank phase 0: vertices per tile
constexpr uint32_t kPrMaxTiles = 1u << 14;      // pagerank: tiles per phase at most
// pagerank iteration tiles: about kPrTargetTiles per phase, of a weight
// (in-edges + 4 x vertices) between kPrWeight and kPrMaxWeight
constexpr uint32_t kPrWeight = 4096;
constexpr uint32_t kPrMaxWeight = 16384;
constexpr uint32_t kPrTargetTiles = 320;
// pagerank per-tile arrays in the slot's kMaxTilesPerReq 8-byte look-back
// words, kPrMaxTiles 4-byte entries each, at these offsets (in entries / kPrMaxTiles)
enum { kPrDang = 0,  // dangling-mass sums, two phases
       kPrL1 = 2,    // L1-change sums
       kPrBv = 3,    // iteration tile j starts in vertex bv[j] ...
       kPrBo = 4,    // ... at in-edge offset bo[j] of it (> 0: the vertex is cut)
       kPrHead = 5,  // sum over the tile's piece of a vertex cut at its start
       kPrTail = 6,  // sum over the tile's piece of a vertex cut at its end
       kPrArrays = 7 };
static_assert(kPrArrays * kPrMaxTiles <= 2 * kMaxTilesPerReq, "pagerank tile arrays exceed the slot area");
constexpr uint32_t kBfsInitTile = 4096;         // bfs: vertices per initialization tile
constexpr uint32_t kBfsTile = 256;              // bfs: frontier vertices per tile (at least; a multiple)
constexpr uint32_t kBfsTargetTiles = 320;       // bfs: tiles per level (about, for big levels)
constexpr uint32_t kMaxPhase = (1u << 24) - 2;  // phases per request (BFS levels + 1, PR iterations + 1)
constexpr uint32_t kQcap = 1u << 20;            // token queue entries
constexpr int32_t kInf = 0x7FFFFFFF;
// Workgroups that read the host ring and the host abort flag.  Host memory is
// read over PCIe and those reads queue behind each other, so only one
// workgroup per XCD polls it; the others see the device words only.
constexpr uint32_t kPollers = 8;
// Entries of the active-request set that idle workgroups scan (one per lane of
// a wave): requests admitted while it is full get help through tokens only.
constexpr uint32_t kActive = 64;

enum { kReqReduce = 0, kReqScan = 1, kReqBfs = 2, kReqPr = 3, kReqStop = 15 };

// request descriptor (host pinned ring; 80 bytes)
struct Req {
  uint32_t seq;        // published when == ring index + 1 (written last by the host)
  uint32_t type;
  uint32_t mem_units;  // admission cost in budget units
  uint32_t arg;        // bfs: source vertex; pagerank: iterations
  uint64_t in;         // reduce/scan input; bfs: CSR row; pr: reverse CSR row
  uint64_t out;        // reduce: int32[1]; scan: uint32[n]; bfs: int32 level[n]; pr: float rank[n]
  uint64_t n;          // elements / vertices
  uint64_t aux0;       // bfs: col; pr: reverse col
  uint64_t aux1;       // bfs: queue scratch (2n u32); pr: scratch (2n floats)
  uint64_t aux2;       // pr: forward CSR row
  float tol;           // pr: stop once the L1 change of an iteration is below (0: run `arg` iterations)
  uint32_t pad[3];
};
static_assert(sizeof(Req) == 80, "request descriptor layout");

// completion record (host pinned; one per ring slot)
struct Done {
  uint32_t done;     // ring index + 1 once complete
  uint32_t status;   // 0 ok
  uint64_t t_admit;  // s_memrealtime ticks (100 MHz) at admission
  uint64_t t_end;    // ticks at completion
  uint32_t phases;   // phases run (pagerank: iterations + 1; bfs: levels + 1)
  uint32_t pad;
};

struct ServeState {  // device memory, 256-byte separated words
  uint32_t claim;  uint32_t p0[63];
  uint32_t qtail;  uint32_t p1[63];  // token queue 0: first phases of admitted requests
  uint32_t qhead;  uint32_t p2[63];
  uint32_t ctail;  uint32_t p7[63];  // token queue 1: later phases (taken first)
  uint32_t chead;  uint32_t p8[63];
  uint32_t stop;   uint32_t p5[63];
  uint32_t aborted; uint32_t p6[63];  // a poller saw the host abort flag (1), or an
                                      // admission wait gave up (5)
  uint32_t budget[kSemWords];         // the memory budget: an EBO priority semaphore
                                      // record of the device API (capacity = units)
  uint32_t p3[kSemWords % 64 ? 64 - kSemWords % 64 : 0];
  uint32_t active[kActive];           // slot + 1 of running requests, 0 = free
};

// Words that every tile updates sit 256 B apart from each other and from the
// fields that tiles only read, so that the atomics of one do not queue behind
// those of another.
struct Slot {  // device copy of an admitted request
  Req r;
  uint32_t ntiles;     // tiles of the current phase
  uint32_t index;
  float base;          // pagerank: (1-a)/n + a*D/n of the current phase's pull
  uint32_t failed;     // a tile's bounded wait gave up
  uint32_t phase;      // current phase
  uint32_t aslot;      // entry in the active-request set, ~0u if none
  uint32_t pr_tiles;   // pagerank: tiles of an iteration phase
  uint64_t pr_w;       // pagerank: weight per iteration tile (in-edges + 4 x vertices)
  uint32_t bfs_ts;     // bfs: frontier vertices per tile of the current level
  alignas(256) uint64_t claim;      // tiles of the current phase << 32 | next tile to hand out
  alignas(256) uint32_t remaining;  // tiles of the current phase not finished
  alignas(256) uint32_t acc;        // reduce: sum of the tile sums
  alignas(256) uint32_t cnt[3];     // bfs: frontier counts, rotated over levels
};

struct ServeArgs {
  Req *ring;
  Done *done;
  volatile uint32_t *abort;
  uint32_t cap;
  ServeState *st;
  Slot *slots;
  uint64_t *tiles;     // token queue 0 entries [kQcap]
  uint64_t *ctiles;    // token queue 1 entries [kQcap]
  uint64_t *lookback;  // [cap][kMaxTilesPerReq]
  uint32_t budget;
  uint32_t spin_limit;
};

__device__ __forceinline__ uint32_t ld_sys(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v) {
  __shared__ uint32_t ws[kWg / 64];
  v = wave_sum(v);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int i = 0; i < kWg / 64; ++i) t += ws[i];
  __syncthreads();
  return t;
}

__device__ __forceinline__ float ld_relaxed_f32(const float *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Agent-coherent stores: visible to every XCD once complete (s_waitcnt), with
// no write-back of the issuing XCD's L2 -- a write-back per tile serializes per
// XCD at microseconds when thousands of small tiles finish per phase.
__device__ __forceinline__ void st_coherent(float *p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// workgroup sum in a fixed order (butterfly per wave, then waves in order)
__device__ __forceinline__ float block_sum_f32(float v) {
  __shared__ float ws[kWg / 64];
  for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.f;
  for (int i = 0; i < kWg / 64; ++i) t += ws[i];
  __syncthreads();
  return t;
}

// the same for doubles
__device__ __forceinline__ double block_sum_f64(double v) {
  __shared__ double ws[kWg / 64];
  for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kWg / 64; ++i) t += ws[i];
  __syncthreads();
  return t;
}

// The memory budget is the device API's EBO priority semaphore
// (hs::sem_wait_master / sem_post_master, the protocol of
// Heterosync/cudaLocksSemaphoreEBO.cu with the priority rule of
// Heterosync/cudaLocksSemaphoreSpin.cu:411-622): an admission takes its request's
// units as one weighted reader, a completion returns them with the poster's
// priority, so waiting admissions back off while requests complete.  The view's
// status word is the loop's abort word: an abort seen by a poller ends the wait,
// and a wait that exceeds the spin limit aborts the loop.
__device__ __forceinline__ ArenaView budget_view(const ServeArgs &p) {
  ArenaView v{};
  v.status = &p.st->aborted;
  v.spin_limit = p.spin_limit;
  return v;
}

__device__ bool budget_acquire(const ServeArgs &p, uint32_t units) {
  return sem_wait_master<true, true>(budget_view(p), p.st->budget, false, units) == kOk;
}

__device__ void budget_release(const ServeArgs &p, uint32_t units) {
  (void)sem_post_master<true>(budget_view(p), p.st->budget, false, units);
}

// ---- tile bodies ----
// reduce tile: 16-byte loads, 8 in flight per thread (a full tile is 64 per
// thread); uint32 wrap, so the order does not matter
__device__ void tile_reduce(Slot &s, uint32_t t) {
  const int32_t *in = (const int32_t *)s.r.in;
  const uint64_t b = (uint64_t)t * kRedTile, e = b + kRedTile < s.r.n ? b + kRedTile : s.r.n;
  uint32_t acc = 0;
  if (e - b == kRedTile && ((uintptr_t)(in + b) & 15u) == 0) {
    const uint4 *p4 = reinterpret_cast<const uint4 *>(in + b) + threadIdx.x;
    for (uint32_t k0 = 0; k0 < kRedTile / 4 / kWg; k0 += 8) {
      uint4 v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = p4[(k0 + k) * kWg];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += v[k].x + v[k].y + v[k].z + v[k].w;
    }
  } else {
    for (uint64_t i = b + threadIdx.x; i < e; i += kWg) acc += (uint32_t)in[i];
  }
  acc = block_sum_u32(acc);
  if (threadIdx.x == 0) add_relaxed(&s.acc, acc);
}

// scan tile: 4 waves x kScanSegs segments of 256 elements; in segment q lane l
// holds elements q*256 + 4l .. 4l+3 of its wave's span (every load and store
// instruction moves 1 KiB contiguous).  The tile's carry comes from a
// wave-parallel decoupled look-back over its predecessors (64 status words per
// step: a ballot finds the nearest inclusive one).
__device__ bool tile_scan(const ServeArgs &p, Slot &s, uint32_t slot, uint32_t t) {
  constexpr uint32_t kSpan = kScanSegs * 256;  // elements per wave
  const uint32_t *in = (const uint32_t *)s.r.in;
  uint32_t *out = (uint32_t *)s.r.out;
  const uint64_t n = s.r.n;
  uint64_t *lb = p.lookback + (size_t)slot * kMaxTilesPerReq;
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint64_t wbase = (uint64_t)t * kScanTile + (uint64_t)wv * kSpan;
  const bool full = wbase + kSpan <= n;
  uint32_t x[kScanSegs][4];
  if (full) {
    const uint4 *src = reinterpret_cast<const uint4 *>(in + wbase) + lane;
#pragma unroll
    for (int q = 0; q < kScanSegs; ++q) {
      const uint4 v = src[q * 64];
      x[q][0] = v.x; x[q][1] = v.y; x[q][2] = v.z; x[q][3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < kScanSegs; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t i = wbase + (uint64_t)q * 256 + 4u * lane + j;
        x[q][j] = i < n ? in[i] : 0u;
      }
  }
  // inclusive scan inside the wave's span; wave total
  uint32_t run = 0;
#pragma unroll
  for (int q = 0; q < kScanSegs; ++q) {
    const uint32_t t4 = x[q][0] + x[q][1] + x[q][2] + x[q][3];
    const uint32_t inc = wave_incl_scan(t4);
    uint32_t pre = run + inc - t4;
    run += __builtin_amdgcn_readlane(inc, 63);
#pragma unroll
    for (int j = 0; j < 4; ++j) { pre += x[q][j]; x[q][j] = pre; }
  }
  __shared__ uint32_t s_wave[kWg / 64];
  __shared__ uint32_t s_prefix;
  __shared__ int s_bad;
  if (lane == 0) s_wave[wv] = run;
  __syncthreads();
  uint32_t before = 0, agg = 0;
#pragma unroll
  for (int k = 0; k < kWg / 64; ++k) {
    before += k < (int)wv ? s_wave[k] : 0u;
    agg += s_wave[k];
  }
  if (wv == 0) {
    uint32_t prefix = 0;
    int bad = 0;
    if (t == 0) {
      if (lane == 0) __hip_atomic_store(lb, (2ull << 32) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(lb + t, (1ull << 32) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int64_t i = (int64_t)t - 1;
      uint32_t polls = 0;
      for (;;) {
        const int64_t idx = i - (int64_t)lane;
        const uint64_t w = idx >= 0 ? __hip_atomic_load(lb + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : (2ull << 32);
        const uint32_t flag = (uint32_t)(w >> 32) & 3u;
        const uint64_t incm = ballot(flag == 2u), readym = ballot(flag != 0u);
        const uint32_t k = incm ? (uint32_t)__builtin_ctzll(incm) : 63u;
        const uint64_t need = k == 63 ? ~0ull : ((2ull << k) - 1ull);
        if ((readym & need) != need) {
          if (++polls > p.spin_limit) { bad = 1; break; }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        prefix += wave_sum(lane <= k ? (uint32_t)w : 0u);
        if (incm) break;
        i -= 64;
      }
      if (lane == 0 && !bad)
        __hip_atomic_store(lb + t, (2ull << 32) | (prefix + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) { s_prefix = prefix; s_bad = bad; }
  }
  __syncthreads();
  if (s_bad) return false;
  const uint32_t add = s_prefix + before;
  if (full) {
    uint4 *dst = reinterpret_cast<uint4 *>(out + wbase) + lane;
#pragma unroll
    for (int q = 0; q < kScanSegs; ++q) dst[q * 64] = make_uint4(x[q][0] + add, x[q][1] + add, x[q][2] + add, x[q][3] + add);
  } else {
#pragma unroll
    for (int q = 0; q < kScanSegs; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t i = wbase + (uint64_t)q * 256 + 4u * lane + j;
        if (i < n) out[i] = x[q][j] + add;
      }
  }
  return true;
}

// BFS phase 0: level initialization of the tile's 4096 vertices
__device__ void tile_bfs_init(Slot &s, uint32_t t) {
  const uint32_t n = (uint32_t)s.r.n, src = s.r.arg;
  int32_t *level = (int32_t *)s.r.out;
  const uint32_t vb = t * kBfsInitTile, ve = min(n, vb + kBfsInitTile);
  for (uint32_t v = vb + threadIdx.x; v < ve; v += kWg) level[v] = v == src ? 0 : kInf;
}

// BFS level L: the tile's 256 frontier vertices, one wave per 64 of them,
// their out-edges 64 at a time (owner lanes by wave_owner), 4 steps of
// candidates in flight, a CAS per newly reached vertex, one queue atomic per
// wave per batch (the reference's visit, LONESTAR/apps/bfs/kernel_gsrb.cu:473-479)
__device__ void tile_bfs_level(Slot &s, uint32_t L, uint32_t t) {
  constexpr int kB = 8;
  const uint32_t n = (uint32_t)s.r.n;
  const uint32_t *row = (const uint32_t *)s.r.in, *col = (const uint32_t *)s.r.aux0;
  int32_t *level = (int32_t *)s.r.out;
  uint32_t *q = (uint32_t *)s.r.aux1;
  const uint32_t nin = ld_relaxed(&s.cnt[L % 3]);
  const uint32_t *qin = q + (size_t)(L & 1) * n;
  uint32_t *qout = q + (size_t)((L + 1) & 1) * n;
  uint32_t *cout = &s.cnt[(L + 1) % 3];
  const uint32_t ts = s.bfs_ts;  // set before the phase opened
  const uint32_t tend = min(nin, (t + 1) * ts);
  for (uint32_t b0 = t * ts + (threadIdx.x >> 6) * 64; b0 < tend; b0 += kWg) {
  const uint32_t i = b0 + lane_id();
  const bool valid = i < tend;
  const uint32_t u = valid ? qin[i] : 0u;
  const uint32_t eb = valid ? row[u] : 0u, deg = valid ? row[u + 1] - eb : 0u;
  const uint32_t incl = wave_incl_scan(deg);
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  const uint32_t excl = incl - deg;
  for (uint32_t off = 0; off < total; off += kB * 64) {
    uint32_t v[kB];
#pragma unroll
    for (int x = 0; x < kB; ++x) {
      const uint32_t e = off + 64 * x + lane_id();
      const bool ev = e < total;
      const uint32_t own = wave_owner(incl, ev ? e : total - 1);
      const uint32_t ob = __shfl(eb, (int)own), oex = __shfl(excl, (int)own);
      v[x] = ev ? col[ob + (e - oex)] : 0xFFFFFFFFu;
    }
    // the filter before the compare-and-swap may read a cached level: the
    // phase began with an acquire (no level older than the previous phase),
    // and a level set during this phase that the cache does not show yet only
    // costs a compare-and-swap that fails
    int32_t lv[kB];
#pragma unroll
    for (int x = 0; x < kB; ++x)
      lv[x] = v[x] != 0xFFFFFFFFu ? __hip_atomic_load(level + v[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
    uint64_t mask[kB];
    uint32_t cnt = 0;
#pragma unroll
    for (int x = 0; x < kB; ++x) {
      bool won = false;
      if (v[x] != 0xFFFFFFFFu && lv[x] == kInf) {
        int32_t expected = kInf;
        won = __hip_atomic_compare_exchange_strong(level + v[x], &expected, (int32_t)L + 1, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      mask[x] = ballot(won);
      cnt += (uint32_t)__builtin_popcountll(mask[x]);
    }
    if (cnt) {
      uint32_t qb = 0;
      if (lane_id() == 0) qb = add_relaxed(cout, cnt);
      qb = __builtin_amdgcn_readfirstlane(qb);
#pragma unroll
      for (int x = 0; x < kB; ++x) {
        if (mask[x] >> lane_id() & 1) qout[qb + lanes_below(mask[x])] = v[x];
        qb += (uint32_t)__builtin_popcountll(mask[x]);
      }
    }
  }
  }  // 64-vertex blocks of the tile
}

// PageRank phase ph of `iters` (normalized pull iterations from x0 = 1/n,
// FP32): for ph >= 1 the tile's vertices pull x_ph = base + a * sum c_{ph-1}
// (blocks of 64 vertices per wave, segmented gather-sums, the order fixed by
// the graph) into out; for ph < iters they then write c_ph = x_ph / deg and the
// tile's dangling mass to part[ph & 1][t].  Contributions are double-buffered
// by phase parity in scratch (2n floats).
// PageRank iteration tiles are ranges of equal weight in the sequence of
// vertices, each weighing 4 plus its in-degree: position F(v) = rrow[v] + 4 v
// starts vertex v, tile j starts at position j wt.  A tile boundary that falls
// among a vertex's in-edges cuts that vertex: the tiles sum their pieces of it
// (head / tail arrays), and the workgroup that closes the phase adds the
// pieces in tile order and finishes the vertex, so a hub's in-edges spread over
// many tiles and the result does not depend on which workgroup ran what.
__device__ void tile_pr(Slot &s, uint32_t ph, uint32_t t, float *part) {
  const uint32_t n = (uint32_t)s.r.n, iters = s.r.arg;
  const uint32_t *rrow = (const uint32_t *)s.r.in, *rcol = (const uint32_t *)s.r.aux0;
  const uint32_t *row = (const uint32_t *)s.r.aux2;
  float *x = (float *)s.r.out;
  float *contrib = (float *)s.r.aux1;
  uint32_t *bv = reinterpret_cast<uint32_t *>(part + kPrBv * kPrMaxTiles);
  uint32_t *bo = reinterpret_cast<uint32_t *>(part + kPrBo * kPrMaxTiles);
  const float a = 0.85f;
  const float *cprev = contrib + (size_t)((ph + 1) & 1) * n;
  float *cnext = contrib + (size_t)(ph & 1) * n;
  const float base = ph ? s.base : 0.f;  // set before the phase opened
  const bool track = ph && s.r.tol > 0.f;  // L1 change of this iteration
  float dl = 0.f, l1 = 0.f;
  if (ph == 0) {
    const uint32_t vb = t * kPrInitTile, ve = min(n, vb + kPrInitTile);
    if (iters) {
      // place the iteration tiles: vertex v holds the starts j wt in [F(v), F(v+1))
      const uint64_t wt = s.pr_w;
      const uint32_t nt = s.pr_tiles;
      for (uint32_t v = vb + threadIdx.x; v < ve; v += kWg) {
        hs_serve::pr_place(v, rrow[v], rrow[v + 1], wt, nt, [&](uint32_t j, uint32_t off) {
          st_coherent(&bv[j], v);
          st_coherent(&bo[j], off);
        });
      }
    }
    for (uint32_t b = vb + (threadIdx.x >> 6) * 64; b < ve; b += kWg) {
      const uint32_t v = b + lane_id();
      if (v >= ve) continue;
      const float xv = 1.0f / (float)n;
      if (iters == 0) st_coherent(&x[v], xv);
      else {
        const uint32_t d = row[v + 1] - row[v];
        st_coherent(&cnext[v], d ? xv / (float)d : 0.f);
        if (!d) dl += xv;
      }
    }
  } else {
    const uint32_t nt = s.pr_tiles;
    const bool last_tile = t + 1 >= nt;
    // starts as phase 0 placed them; the clamps keep every index in range
    const hs_serve::PrTileSpan sp =
        hs_serve::pr_tile_span(t, nt, n, t ? bv[t] : 0u, t ? bo[t] : 0u, last_tile ? 0u : bv[t + 1],
                               last_tile ? 0u : bo[t + 1]);
    float *head = part + kPrHead * kPrMaxTiles, *tail = part + kPrTail * kPrMaxTiles;
    for (uint32_t b = sp.v0 + (threadIdx.x >> 6) * 64; b < sp.vend; b += kWg) {
      const uint32_t v = b + lane_id();
      const bool valid = v < sp.vend;
      const uint32_t rb = valid ? rrow[v] : 0u, re = valid ? rrow[v + 1] : 0u;
      const hs_serve::PrPiece pc = hs_serve::pr_piece(sp, v, rb, re);
      const bool cut_head = valid && pc.cut_head, cut_tail = valid && pc.cut_tail;
      const uint32_t es = pc.es, ee = pc.ee;
      const float sum = wave_run_gather_sum<8>(rcol, cprev, es, ee - es);
      if (cut_head) st_coherent(&head[t], sum);  // also a tile inside one vertex
      else if (cut_tail) st_coherent(&tail[t], sum);
      else if (valid) {
        const float xv = base + a * sum;
        if (track) l1 += fabsf(xv - (ph == 1 ? 1.0f / (float)n : x[v]));
        st_coherent(&x[v], xv);
        if (ph < iters) {
          const uint32_t d = row[v + 1] - row[v];
          st_coherent(&cnext[v], d ? xv / (float)d : 0.f);
          if (!d) dl += xv;
        }
      }
    }
  }
  if (ph < iters) {
    const float tot = block_sum_f32(dl);
    if (threadIdx.x == 0) st_coherent(&part[(size_t)(kPrDang + (ph & 1)) * kPrMaxTiles + t], tot);
  }
  if (track) {
    const float tot = block_sum_f32(l1);
    if (threadIdx.x == 0) st_coherent(&part[(size_t)kPrL1 * kPrMaxTiles + t], tot);
  }
}

// End of PageRank phase ph >= 1 (the closing workgroup): finishes the cut
// vertices -- boundary j with bo[j] > 0 cuts bv[j]; the first boundary of a
// cut vertex adds its pieces (the tail of tile j-1, then the heads of the tiles
// it starts) -- and returns their dangling mass and L1 change (block sums).
__device__ void pr_finish_cut(Slot &s, uint32_t ph, uint32_t ntl, float *part, float &cdl, float &cl1) {
  const uint32_t n = (uint32_t)s.r.n, iters = s.r.arg;
  const uint32_t *row = (const uint32_t *)s.r.aux2;
  float *x = (float *)s.r.out;
  float *cnext = (float *)s.r.aux1 + (size_t)(ph & 1) * n;
  const uint32_t *bv = reinterpret_cast<const uint32_t *>(part + kPrBv * kPrMaxTiles);
  const uint32_t *bo = reinterpret_cast<const uint32_t *>(part + kPrBo * kPrMaxTiles);
  const float *head = part + kPrHead * kPrMaxTiles, *tail = part + kPrTail * kPrMaxTiles;
  const float base = ld_relaxed_f32(&s.base), a = 0.85f;
  const bool track = s.r.tol > 0.f;
  float dl = 0.f, l1 = 0.f;
  for (uint32_t j = 1 + threadIdx.x; j < ntl; j += kWg) {
    const uint32_t u = min(bv[j], n - 1);
    if (!bo[j] || (j > 1 && bv[j - 1] == bv[j] && bo[j - 1])) continue;
    float sum = tail[j - 1];
    for (uint32_t k = j; k < ntl && bv[k] == bv[j] && bo[k]; ++k) sum += head[k];
    const float xv = base + a * sum;
    if (track) l1 += fabsf(xv - (ph == 1 ? 1.0f / (float)n : x[u]));
    x[u] = xv;
    if (ph < iters) {
      const uint32_t d = row[u + 1] - row[u];
      cnext[u] = d ? xv / (float)d : 0.f;
      if (!d) dl += xv;
    }
  }
  cdl = block_sum_f32(dl);
  cl1 = block_sum_f32(l1);
  wait_stores();
  __syncthreads();
}

// tiles of a request's first phase
__device__ uint32_t ntiles_of(const Req &r) {
  switch (r.type) {
    case kReqReduce: return r.n ? (uint32_t)((r.n + kRedTile - 1) / kRedTile) : 1u;
    case kReqScan: return r.n ? (uint32_t)((r.n + kScanTile - 1) / kScanTile) : 1u;
    case kReqBfs: return (uint32_t)((r.n + kBfsInitTile - 1) / kBfsInitTile);
    case kReqPr: return (uint32_t)((r.n + kPrInitTile - 1) / kPrInitTile);
    default: return 1;
  }
}

// Work of a phase is handed out on two levels.  The slot's claim word holds
// the phase's tile count (high half) and the next tile to hand out (low half):
// a workgroup working on the request takes tiles with a fetch-and-add on it
// (in increasing order, which the scan's look-back relies on) until the count
// is reached.  A phase is opened by the workgroup that finished the previous
// one (or admitted the request), which then works on it itself; tokens in the
// tile queues -- tickets to join the slot's request -- bring more workgroups,
// about 2 sqrt(tiles) of them, so that claiming a token from a shared queue
// head (a compare-and-swap that every idle workgroup contends for) is paid a
// few times per phase instead of once per tile.
// A tile taken is always one of the current phase: the phase cannot end while
// a tile of it is taken and unfinished, the fields of the next phase are
// written before its claim word, and a late fetch-and-add on a finished
// phase's word (whose count is reached) only raises the low half, which the
// next phase's word replaces.  A token that arrives late joins whatever phase
// or request the slot then holds, or leaves at once.

// token entry (+1, so 0 means empty): the slot index
__device__ __forceinline__ uint64_t token_entry(uint32_t slot) { return (uint64_t)slot + 1; }

__device__ __forceinline__ uint32_t extra_tokens(uint32_t tiles) {
  uint32_t r = (uint32_t)__builtin_sqrtf((float)tiles);
  while (r * r < tiles) ++r;
  const uint32_t k = 2 * r;
  return tiles ? (k < tiles ? k : tiles) - 1 : 0u;  // the opener is one of them
}

// open phase `phase` of `tiles` tiles of a slot's request (master thread):
// the phase's fields, then the claim word, then the tokens (one more when the
// opener does not join the phase).  False if the token queue stayed full for
// the spin limit.
// A request's first phase sends its tokens to queue 0, behind what was
// admitted before it; later phases use queue 1, which workgroups drain first,
// so a request that has started is not sent back behind every request
// admitted meanwhile.
__device__ bool open_phase(const ServeArgs &p, uint32_t slot, uint32_t phase, uint32_t tiles, bool joins) {
  Slot &s = p.slots[slot];
  const uint32_t count = extra_tokens(tiles) + (joins || !tiles ? 0u : 1u);
  uint32_t *tail = phase ? &p.st->ctail : &p.st->qtail;
  uint32_t *head = phase ? &p.st->chead : &p.st->qhead;
  uint64_t *q = phase ? p.ctiles : p.tiles;
  uint32_t polls = 0;
  while (ld_relaxed(tail) - ld_relaxed(head) + count > kQcap) {
    if (++polls > p.spin_limit) return false;
    __builtin_amdgcn_s_sleep(4);
  }
  st_relaxed(&s.phase, phase);
  st_relaxed(&s.ntiles, tiles);
  st_relaxed(&s.remaining, tiles);
  fence_release();
  __hip_atomic_store(&s.claim, (uint64_t)tiles << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (count) {
    const uint32_t at = add_relaxed(tail, count);
    for (uint32_t i = 0; i < count; ++i)
      __hip_atomic_store(q + ((at + i) % kQcap), token_entry(slot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// claim the head entry of a token queue (master thread): 1 and *entry = the
// entry, 0 if the queue is empty or another workgroup won the head, -1 if the
// claimed entry was never published within the spin limit
__device__ int take_token(const ServeArgs &p, uint32_t *headp, const uint32_t *tailp, uint64_t *q,
                          uint64_t *entry) {
  const uint32_t head = ld_relaxed(headp);
  if (head >= ld_relaxed(tailp)) return 0;
  uint32_t h = head;
  if (!__hip_atomic_compare_exchange_strong(headp, &h, head + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT))
    return 0;
  uint64_t *ent = q + (head % kQcap);
  uint64_t e;
  uint32_t polls = 0;
  while ((e = __hip_atomic_load(ent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
    if (++polls > p.spin_limit) return -1;
    __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(ent, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  *entry = e;
  return 1;
}

__global__ void __launch_bounds__(kWg) k_serve(ServeArgs p) {
  __shared__ uint64_t s_task;  // 0 none; 1<<62 | slot : token; 2<<62 | index : admission
  __shared__ int s_exit, s_ok;
  __shared__ uint32_t s_t, s_ph, s_type, s_help;
  __shared__ int s_last, s_more;
  uint32_t idle = 0, iter = 0;
  uint64_t idle_units = 0;  // s_sleep units slept since the last work
  const bool poller = blockIdx.x < kPollers;
  // (slot, request index, phase) of this workgroup's last tile (master thread):
  // a tile of the same phase needs no acquire, the earlier one covered it
  uint32_t seen_slot = ~0u, seen_index = 0, seen_phase = 0;
  for (;;) {
    ++iter;
    if (threadIdx.x == 0) {
      s_task = 0;
      s_exit = 0;
      // pollers look at the host ring first, so admission never waits behind tiles
      if (poller) {
        const uint32_t c = ld_relaxed(&p.st->claim);
        const Req *r = p.ring + (c % p.cap);
        if (ld_sys(&r->seq) == c + 1) {
          uint32_t cc = c;
          if (__hip_atomic_compare_exchange_strong(&p.st->claim, &cc, c + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
            s_task = (2ull << 62) | c;
        }
      }
      uint64_t e = 0;
      int got = 0;
      if (!s_task) got = take_token(p, &p.st->chead, &p.st->ctail, p.ctiles, &e);
      if (!s_task && got == 0) got = take_token(p, &p.st->qhead, &p.st->qtail, p.tiles, &e);
      if (got > 0) s_task = (1ull << 62) | (e - 1);
      if (got < 0) s_exit = 2;
      if (!s_task && !s_exit && ld_relaxed(&p.st->stop) && ld_relaxed(&p.st->qhead) >= ld_relaxed(&p.st->qtail) &&
          ld_relaxed(&p.st->chead) >= ld_relaxed(&p.st->ctail))
        s_exit = 1;
      if (poller && (iter & 63u) == 0 && ld_sys((const uint32_t *)p.abort)) st_relaxed(&p.st->aborted, 1u);
      if (ld_relaxed(&p.st->aborted)) s_exit = 3;
    }
    __syncthreads();
    const uint64_t task = s_task;
    const int ex = s_exit;
    __syncthreads();
    if (ex) return;
    if (task) {
      idle = 0;
      idle_units = 0;
    }
    uint32_t slot;
    if (!task) {
      // no token and no admission: join a running request that has tiles
      // left, scanning the active set from a workgroup-dependent entry so
      // that idle workgroups spread over the running requests
      if (threadIdx.x < 64) {
        const uint32_t e = ld_relaxed(&p.st->active[(threadIdx.x + blockIdx.x) % kActive]);
        bool avail = false;
        if (e) {
          const uint64_t c = __hip_atomic_load(&p.slots[e - 1].claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          avail = (uint32_t)c < (uint32_t)(c >> 32);
        }
        const uint64_t m = __ballot(avail);
        const uint32_t h = __shfl(e, m ? __ffsll((unsigned long long)m) - 1 : 0);
        if (threadIdx.x == 0) s_help = m ? h : 0u;
      }
      __syncthreads();
      const uint32_t h = s_help;
      __syncthreads();
      if (!h) {
        // back off while idle (the polls of hundreds of idle workgroups slow
        // the atomics of the busy ones); pollers stay quicker for the ring.
        // The loop ends after as long idle as spin_limit polls of 8 units.
        const uint32_t units = idle < 64 ? 8u : (poller ? 32u : 128u);
        ++idle;
        idle_units += units;
        if (idle_units > 8ull * p.spin_limit) return;
        sleep_units(units);
        continue;
      }
      slot = h - 1;
    } else if ((task >> 62) == 2) {
      // admission of request index c into slot c % cap; this workgroup then
      // works on the request's first phase
      const uint32_t c = (uint32_t)task;
      slot = c % p.cap;
      Slot &s = p.slots[slot];
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope: the host's descriptor
        const Req *hr = p.ring + slot;
        s.r = *hr;
        if (s.r.type == kReqStop) {
          st_relaxed(&p.st->stop, 1u);
          __hip_atomic_store(&p.done[slot].done, c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
          s_ok = 0;
        } else if (!budget_acquire(p, s.r.mem_units)) {
          st_relaxed(&p.st->stop, 1u);
          s_ok = -1;
        } else {
          p.done[slot].t_admit = wall_clock64();
          st_relaxed(&s.index, c);
          s.acc = 0;
          s.failed = 0;
          s.base = 0.f;
          s.cnt[0] = 1;  // bfs: the level-0 frontier is the source
          s.cnt[1] = s.cnt[2] = 0;
          if (s.r.type == kReqBfs) ((uint32_t *)s.r.aux1)[0] = s.r.arg;
          if (s.r.type == kReqPr) {
            // iteration tiles: weight w(v) = rrow[v] + 4 v up to v, total W = w(n);
            // about kPrTargetTiles tiles, weight clamped to [kPrWeight,
            // kPrMaxWeight], at most kPrMaxTiles tiles, and (tiles - 1) wt < W.
            // Measured (20 iterations, 512 workgroups): R-MAT 16 best at 4096,
            // R-MAT 18 and 20 at 16384 (R-MAT 20: 25 ms at 4096, 9.2 ms at
            // 16384, 11.4 ms at 55K)
            const uint64_t W = (uint64_t)((const uint32_t *)s.r.in)[s.r.n] + 4ull * s.r.n;
            uint64_t wt = (W + kPrTargetTiles - 1) / kPrTargetTiles;
            wt = wt < kPrMaxWeight ? wt : kPrMaxWeight;
            wt = wt > kPrWeight ? wt : kPrWeight;
            const uint64_t wmin = (W + kPrMaxTiles - 1) / kPrMaxTiles;
            wt = wt > wmin ? wt : wmin;
            s.pr_w = wt;
            s.pr_tiles = (uint32_t)((W + wt - 1) / wt);
          }
          s_ok = 1;
        }
      }
      __syncthreads();
      if (s_ok < 0) return;
      if (s_ok == 0) continue;
      const uint32_t nt = ntiles_of(s.r);
      if (s.r.type == kReqScan) {
        uint64_t *lb = p.lookback + (size_t)slot * kMaxTilesPerReq;
        for (uint32_t i = threadIdx.x; i < nt; i += kWg) lb[i] = 0;
      }
      wait_stores();
      __syncthreads();
      if (threadIdx.x == 0) {
        // listed before its first phase opens, so listed before it can complete
        s.aslot = ~0u;
        for (uint32_t i = 0; i < kActive; ++i) {
          const uint32_t at = (c + i) % kActive;
          uint32_t z = 0;
          if (ld_relaxed(&p.st->active[at]) == 0u &&
              __hip_atomic_compare_exchange_strong(&p.st->active[at], &z, slot + 1, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            s.aslot = at;
            break;
          }
        }
        if (!open_phase(p, slot, 0, nt, !poller)) s_ok = -1;
      }
      __syncthreads();
      if (s_ok < 0) return;
      // a poller goes back to the host ring: admission must not wait until
      // the requests it admitted are done
      if (poller) continue;
    } else {
      slot = (uint32_t)(task & 0xFFFFu);
    }
    idle = 0;
    idle_units = 0;
    // work on the slot's request: take tiles of its current phase until the
    // claim word says none is left
    Slot &s = p.slots[slot];
    float *part = reinterpret_cast<float *>(p.lookback + (size_t)slot * kMaxTilesPerReq);
    // Finished tiles are counted into the phase's arrival counter together,
    // once the claim word has no tile left for this workgroup: one release
    // and one atomic per workgroup and phase instead of per tile.  The phase
    // cannot end before the batch is counted, so every claim of a batch is of
    // one phase of one request.
    uint32_t nb = 0, bph = 0, btype = 0;
    for (;;) {
      if (threadIdx.x == 0) {
        uint32_t t = ~0u;
        if (ld_relaxed(&p.st->aborted)) {
          t = ~1u;
        } else {
          const uint64_t c = __hip_atomic_fetch_add(&s.claim, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((uint32_t)c < (uint32_t)(c >> 32)) {
            t = (uint32_t)c;
            if (!nb) {  // later tiles of a batch are of the same phase: s_ph, s_type hold
              __atomic_signal_fence(__ATOMIC_SEQ_CST);
              const uint32_t ph = ld_relaxed(&s.phase), ix = ld_relaxed(&s.index);
              if (ph != seen_phase || ix != seen_index || slot != seen_slot) {
                // first tile of this phase here: the phase's fields and what
                // earlier phases wrote (released before the claim word)
                fence_acquire();
                seen_slot = slot;
                seen_index = ix;
                seen_phase = ph;
              }
              s_ph = ph;
              s_type = ld_relaxed(&s.r.type);
            }
          }
        }
        s_t = t;
      }
      __syncthreads();
      const uint32_t t = s_t;
      __syncthreads();
      if (t == ~1u) return;
      if (t != ~0u) {
        const uint32_t tph = s_ph, ttype = s_type;
        bool ok = true;
        if (ttype == kReqReduce) tile_reduce(s, t);
        else if (ttype == kReqScan) ok = tile_scan(p, s, slot, t);
        else if (ttype == kReqBfs) { if (tph == 0) tile_bfs_init(s, t); else tile_bfs_level(s, tph - 1, t); }
        else if (ttype == kReqPr) tile_pr(s, tph, t, part);
        wait_stores();
        __syncthreads();
        if (threadIdx.x == 0 && !ok) st_relaxed(&s.failed, 1u);  // reported with the completion
        ++nb;
        bph = tph;
        btype = ttype;
        continue;  // next tile of the phase
      }
      if (!nb) break;
      const uint32_t ph = bph, type = btype;
      if (threadIdx.x == 0) {
        // a PageRank tile wrote everything with agent-coherent stores, which
        // every wave has waited for, and a reduce tile only adds atomically:
        // no L2 write-back needed
        if (type != kReqPr && type != kReqReduce) fence_release();
        s_last = add_relaxed(&s.remaining, 0u - nb) == nb;  // the phase's last tiles: fan-in
        if (s_last) fence_acquire();                         // what every tile of the phase wrote
      }
      nb = 0;
      __syncthreads();
      if (!s_last) break;
      // the phase is complete: open the next one, or complete the request
      uint32_t next = 0;
      const uint32_t ntl = ld_relaxed(&s.ntiles);
      bool converged = false;
      float cdl = 0.f, cl1 = 0.f;  // of the cut vertices
      if (type == kReqPr && ph) pr_finish_cut(s, ph, ntl, part, cdl, cl1);
      if (type == kReqPr && ph && s.r.tol > 0.f) {
        double c = 0.0;  // L1 change of iteration ph, tiles in order
        for (uint32_t i = threadIdx.x; i < ntl; i += kWg) c += ld_relaxed_f32(part + (size_t)kPrL1 * kPrMaxTiles + i);
        converged = block_sum_f64(c) + (double)cl1 < (double)s.r.tol;
      }
      if (type == kReqPr && ph < s.r.arg && !converged) {
        float dl = 0.f;  // dangling mass of x_ph, tiles in order
        for (uint32_t i = threadIdx.x; i < ntl; i += kWg)
          dl += ld_relaxed_f32(part + (size_t)(kPrDang + (ph & 1)) * kPrMaxTiles + i);
        const float D = block_sum_f32(dl) + cdl;
        const float a = 0.85f, nf = (float)s.r.n;
        if (threadIdx.x == 0) __hip_atomic_store(&s.base, (1.f - a) / nf + a * D / nf, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        next = ld_relaxed(&s.pr_tiles);
      } else if (type == kReqBfs) {
        const uint32_t nc = ph == 0 ? 1u : ld_relaxed(&s.cnt[ph % 3]);  // level ph-1 wrote cnt[ph % 3]
        // about kBfsTargetTiles tiles of a multiple of kBfsTile vertices
        uint32_t ts = (nc + kBfsTargetTiles - 1) / kBfsTargetTiles;
        ts = ts > kBfsTile ? (ts + kBfsTile - 1) / kBfsTile * kBfsTile : kBfsTile;
        next = (nc + ts - 1) / ts;
        if (threadIdx.x == 0) st_relaxed(&s.bfs_ts, ts);
      }
      if (threadIdx.x == 0) {
        s_more = 0;
        bool done_ok = true;
        if (next) {
          if (type == kReqBfs && ph > 0) st_relaxed(&s.cnt[(ph + 1) % 3], 0u);  // level ph's output count
          if (ph + 1 <= kMaxPhase && open_phase(p, slot, ph + 1, next, !poller)) s_more = 1;
          else done_ok = false;
        }
        if (!s_more) {  // completion record
          if (ld_relaxed(&s.failed)) done_ok = false;
          if (type == kReqReduce) ((int32_t *)s.r.out)[0] = (int32_t)ld_relaxed(&s.acc);
          wait_stores();
          fence_release();
          Done *d = p.done + slot;
          d->status = done_ok ? 0u : 5u;
          d->phases = ph + 1;
          d->t_end = wall_clock64();
          budget_release(p, s.r.mem_units);
          const uint32_t as = ld_relaxed(&s.aslot);
          if (as != ~0u) st_relaxed(&p.st->active[as], 0u);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
          __hip_atomic_store(&d->done, s.index + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
      __syncthreads();
      if (poller) break;  // the next phase has a token for it
      // after a completion the claim word's count is reached: the loop ends
    }
  }
}

struct Server {
  Req *ring = nullptr;
  Done *done = nullptr;
  uint32_t *abort = nullptr;
  void *dev = nullptr;
  ServeArgs args{};
  hipStream_t stream = nullptr;
  uint32_t next = 0;
};

}  // namespace

extern "C" {

typedef struct hs_server hs_server;

int hs_serve_start(int num_wgs, uint32_t ring_capacity, uint32_t budget_units, hs_server **out) {
  if (!out || ring_capacity == 0 || ring_capacity > 4096 || budget_units == 0 || budget_units == ~0u)
    return HS_ERR_INVALID;
  const void *k = (const void *)k_serve;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 1);
  if (hs_host::check_residency(k, num_wgs, kWg, 0) != HS_OK) return HS_ERR_RESIDENCY;
  Server *s = new Server();
  // releases what a failed start allocated (their own failures add nothing)
  auto fail = [s](int rc) {
    if (s->dev) (void)hipFree(s->dev);
    if (s->abort) (void)hipHostFree(s->abort);
    if (s->done) (void)hipHostFree(s->done);
    if (s->ring) (void)hipHostFree(s->ring);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return rc;
  };
  // a stream of its own that does not synchronize with the null stream, so
  // the host can copy results while the loop runs (default priority)
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return fail(HS_ERR_HIP);
  if (hipHostMalloc((void **)&s->ring, sizeof(Req) * ring_capacity, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&s->done, sizeof(Done) * ring_capacity, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&s->abort, 256, hipHostMallocCoherent) != hipSuccess)
    return fail(HS_ERR_NOMEM);
  memset(s->ring, 0, sizeof(Req) * ring_capacity);
  memset(s->done, 0, sizeof(Done) * ring_capacity);
  memset(s->abort, 0, 256);
  const size_t head = sizeof(ServeState) + sizeof(Slot) * ring_capacity + 2 * sizeof(uint64_t) * kQcap;
  const size_t bytes = head + sizeof(uint64_t) * (size_t)ring_capacity * kMaxTilesPerReq;
  if (hipMalloc(&s->dev, bytes) != hipSuccess) return fail(HS_ERR_NOMEM);
  std::vector<char> zero(head, 0);
  ServeState init{};
  init.budget[kSemCount] = init.budget[kSemMax] = budget_units + 1u;  // capacity + 1, as hs_sem_create
  memcpy(zero.data(), &init, sizeof(init));
  if (hipMemcpyAsync(s->dev, zero.data(), head, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return fail(HS_ERR_HIP);
  char *w = (char *)s->dev;
  ServeArgs &a = s->args;
  a.st = (ServeState *)w; w += sizeof(ServeState);
  a.slots = (Slot *)w; w += sizeof(Slot) * ring_capacity;
  a.tiles = (uint64_t *)w; w += sizeof(uint64_t) * kQcap;
  a.ctiles = (uint64_t *)w; w += sizeof(uint64_t) * kQcap;
  a.lookback = (uint64_t *)w;
  a.ring = s->ring;
  a.done = s->done;
  a.abort = s->abort;
  a.cap = ring_capacity;
  a.budget = budget_units;
  a.spin_limit = 1u << 28;
  void *args[] = {(void *)&s->args};
  if (hipLaunchKernel(k, dim3(num_wgs), dim3(kWg), args, 0, s->stream) != hipSuccess) return fail(HS_ERR_HIP);
  *out = (hs_server *)s;
  return HS_OK;
}

}  // extern "C"

namespace {
int submit(hs_server *srv, uint32_t type, uint32_t mem_units, uint32_t arg, float tol, uint64_t in, uint64_t out,
           uint64_t n, uint64_t aux0, uint64_t aux1, uint64_t aux2, uint32_t *index_out) {
  Server *s = (Server *)srv;
  if (!s || !index_out) return HS_ERR_INVALID;
  if (type != kReqReduce && type != kReqScan && type != kReqBfs && type != kReqPr && type != kReqStop)
    return HS_ERR_INVALID;
  if (type == kReqScan && (n + kScanTile - 1) / kScanTile > kMaxTilesPerReq) return HS_ERR_INVALID;
  if ((type == kReqBfs || type == kReqPr) && (n == 0 || n > 0xFFFFFFFFull)) return HS_ERR_INVALID;
  // bfs: 2^28 vertices at most; pagerank: at most kPrMaxTiles tiles per
  // phase, whose sums and starts share the slot's look-back words
  if (type == kReqBfs && n > (1ull << 28)) return HS_ERR_INVALID;
  if (type == kReqPr && ((n + kPrInitTile - 1) / kPrInitTile > kPrMaxTiles || arg >= kMaxPhase)) return HS_ERR_INVALID;
  if (type == kReqBfs && arg >= n) return HS_ERR_INVALID;
  if (mem_units > s->args.budget) return HS_ERR_INVALID;
  const uint32_t idx = s->next;
  const uint32_t slot = idx % s->args.cap;
  // a slot is reusable once its previous request completed
  if (idx >= s->args.cap && __atomic_load_n(&s->done[slot].done, __ATOMIC_ACQUIRE) != idx - s->args.cap + 1)
    return HS_ERR_NOMEM;
  Req *r = s->ring + slot;
  r->type = type;
  r->mem_units = mem_units;
  r->arg = arg;
  r->in = in;
  r->out = out;
  r->n = n;
  r->aux0 = aux0;
  r->aux1 = aux1;
  r->aux2 = aux2;
  r->tol = tol;
  __atomic_store_n(&s->done[slot].done, 0u, __ATOMIC_RELAXED);
  __atomic_store_n(&r->seq, idx + 1, __ATOMIC_RELEASE);
  s->next = idx + 1;
  *index_out = idx;
  return HS_OK;
}
}  // namespace

extern "C" {

int hs_serve_submit(hs_server *srv, uint32_t type, uint32_t mem_units, uint32_t arg, uint64_t in, uint64_t out,
                    uint64_t n, uint64_t aux0, uint64_t aux1, uint64_t aux2, uint32_t *index_out) {
  return submit(srv, type, mem_units, arg, 0.f, in, out, n, aux0, aux1, aux2, index_out);
}

int hs_serve_submit_pagerank(hs_server *srv, uint32_t mem_units, uint32_t max_iter, float tol, uint64_t rrow,
                             uint64_t rank, uint64_t n, uint64_t rcol, uint64_t scratch, uint64_t row,
                             uint32_t *index_out) {
  if (!(tol >= 0.f) || tol > 1e30f) return HS_ERR_INVALID;  // rejects NaN and infinity
  return submit(srv, kReqPr, mem_units, max_iter, tol, rrow, rank, n, rcol, scratch, row, index_out);
}

// 1 when request `index` completed (status / device ticks filled), 0 if not yet
int hs_serve_poll(hs_server *srv, uint32_t index, uint32_t *status, uint64_t *t_admit, uint64_t *t_end) {
  Server *s = (Server *)srv;
  if (!s) return -HS_ERR_INVALID;
  const Done *d = s->done + (index % s->args.cap);
  if (__atomic_load_n(&d->done, __ATOMIC_ACQUIRE) != index + 1) return 0;
  if (status) *status = d->status;
  if (t_admit) *t_admit = d->t_admit;
  if (t_end) *t_end = d->t_end;
  return 1;
}

int hs_serve_phases(hs_server *srv, uint32_t index, uint32_t *phases) {
  Server *s = (Server *)srv;
  if (!s || !phases) return HS_ERR_INVALID;
  const Done *d = s->done + (index % s->args.cap);
  if (__atomic_load_n(&d->done, __ATOMIC_ACQUIRE) != index + 1) return HS_ERR_INVALID;
  *phases = d->phases;
  return HS_OK;
}

// stop the loop (after the queued requests unless abort) and free it
int hs_serve_stop(hs_server *srv, int abort) {
  Server *s = (Server *)srv;
  if (!s) return HS_ERR_INVALID;
  if (abort) {
    __atomic_store_n(s->abort, 1u, __ATOMIC_SEQ_CST);
  } else {
    uint32_t idx;
    int rc, tries = 0;
    // the ring slot frees when its previous request completes; give up after ~10 s
    while ((rc = hs_serve_submit(srv, kReqStop, 0, 0, 0, 0, 0, 0, 0, 0, &idx)) == HS_ERR_NOMEM && ++tries < 100000)
      usleep(100);
    if (rc) __atomic_store_n(s->abort, 1u, __ATOMIC_SEQ_CST);
  }
  hipError_t e = hipStreamSynchronize(s->stream);
  // release everything; report the first failure
  const hipError_t r[5] = {hipStreamDestroy(s->stream), hipFree(s->dev), hipHostFree(s->ring),
                           hipHostFree(s->done), hipHostFree(s->abort)};
  for (hipError_t x : r)
    if (e == hipSuccess) e = x;
  delete s;
  return e == hipSuccess ? HS_OK : HS_ERR_HIP;
}

}  // extern "C"
