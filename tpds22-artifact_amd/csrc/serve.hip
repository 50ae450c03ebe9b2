// serve.hip -- a serving loop that runs on the device as one persistent kernel.
//
// No reference counterpart.  The host publishes requests into a ring of
// descriptors in coherent pinned host memory; every workgroup of the
// persistent grid (1) takes the next ready tile of any admitted request, or
// (2) claims the next published request, admits it when its working-memory
// cost fits the budget -- a counting semaphore whose releases (request
// completions) raise its priority word so that waiting admissions back off, the
// priority rule of Heterosync/cudaLocksSemaphoreSpin.cu:411-622 -- and appends
// its tiles to a shared tile queue.  Tiles of different requests interleave in
// that queue, so requests that arrive while others run are admitted and
// overlap on the device (continuous batching).  The last tile of a request
// writes its completion record to host memory.
//   reduce_i32 : tiles of 64Ki elements, sums combined with atomics (wrap)
//   scan_u32   : tiles of 4096 elements, carries by decoupled look-back
//   bfs        : one tile; a workgroup-level queue BFS (small graphs)
//   pagerank   : one tile; a workgroup-level pull power iteration (small graphs)
// The kernel runs on its own non-blocking stream, every wait is bounded, and an
// abort word in host memory stops the loop.
#include <hip/hip_runtime.h>

#include <string.h>
#include <unistd.h>

#include <vector>

#include "hs_internal.hpp"
#include "wave.hpp"

using namespace hs;

namespace {

constexpr int kWg = 256;
constexpr uint32_t kScanTile = 4096;
constexpr uint32_t kRedTile = 65536;
constexpr uint32_t kMaxTilesPerReq = 1u << 16;  // scan: 2^28 elements per request at most
constexpr uint32_t kQcap = 1u << 20;            // tile queue entries
constexpr int32_t kInf = 0x7FFFFFFF;
// Workgroups that read the host ring and the host abort flag.  Host memory is
// read over PCIe and those reads queue behind each other, so only one
// workgroup per XCD polls it; the others see the device words only.
constexpr uint32_t kPollers = 8;

enum { kReqReduce = 0, kReqScan = 1, kReqBfs = 2, kReqPr = 3, kReqStop = 15 };

// request descriptor (host pinned ring; 64 bytes)
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
};

// completion record (host pinned; one per ring slot)
struct Done {
  uint32_t done;     // ring index + 1 once complete
  uint32_t status;   // 0 ok
  uint64_t t_admit;  // s_memrealtime ticks (100 MHz) at admission
  uint64_t t_end;    // ticks at completion
  uint64_t pad;
};

struct ServeState {  // device memory, 256-byte separated words
  uint32_t claim;  uint32_t p0[63];
  uint32_t qtail;  uint32_t p1[63];
  uint32_t qhead;  uint32_t p2[63];
  uint32_t avail;  uint32_t p3[63];
  uint32_t prio;   uint32_t p4[63];
  uint32_t stop;   uint32_t p5[63];
  uint32_t aborted; uint32_t p6[63];  // a poller saw the host abort flag
};

struct Slot {  // device copy of an admitted request
  Req r;
  uint32_t remaining;
  uint32_t acc;
  uint32_t ntiles;
  uint32_t index;
};

struct ServeArgs {
  Req *ring;
  Done *done;
  volatile uint32_t *abort;
  uint32_t cap;
  ServeState *st;
  Slot *slots;
  uint64_t *tiles;
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

__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t &total) {
  __shared__ uint32_t ws[kWg / 64];
  const uint32_t inc = wave_incl_scan(v);
  if (lane_id() == 63) ws[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint32_t before = 0, t = 0;
  for (int i = 0; i < kWg / 64; ++i) {
    before += i < (int)(threadIdx.x >> 6) ? ws[i] : 0u;
    t += ws[i];
  }
  __syncthreads();
  total = t;
  return before + inc - v;
}

// budget: counting semaphore, releases have priority (master thread)
__device__ bool budget_acquire(const ServeArgs &p, uint32_t units) {
  uint32_t polls = 0;
  for (;;) {
    while (ld_relaxed(&p.st->prio) != 0u) {
      if (++polls > p.spin_limit) return false;
      __builtin_amdgcn_s_sleep(2);
    }
    uint32_t cur = ld_relaxed(&p.st->avail);
    while (cur >= units) {
      if (__hip_atomic_compare_exchange_strong(&p.st->avail, &cur, cur - units, __ATOMIC_ACQUIRE,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return true;
    }
    if (++polls > p.spin_limit || ((polls & 255u) == 0 && ld_relaxed(&p.st->aborted))) return false;
    __builtin_amdgcn_s_sleep(2);
  }
}

__device__ void budget_release(const ServeArgs &p, uint32_t units) {
  add_relaxed(&p.st->prio, 1u);
  add_release(&p.st->avail, units);
  sub_release(&p.st->prio, 1u);
}

// ---- tile bodies ----
__device__ void tile_reduce(Slot &s, uint32_t t) {
  const int32_t *in = (const int32_t *)s.r.in;
  const uint64_t b = (uint64_t)t * kRedTile, e = b + kRedTile < s.r.n ? b + kRedTile : s.r.n;
  uint32_t acc = 0;
  for (uint64_t i = b + threadIdx.x; i < e; i += kWg) acc += (uint32_t)in[i];
  acc = block_sum_u32(acc);
  if (threadIdx.x == 0) add_relaxed(&s.acc, acc);
}

__device__ bool tile_scan(const ServeArgs &p, Slot &s, uint32_t slot, uint32_t t) {
  const uint32_t *in = (const uint32_t *)s.r.in;
  uint32_t *out = (uint32_t *)s.r.out;
  const uint64_t n = s.r.n;
  uint64_t *lb = p.lookback + (size_t)slot * kMaxTilesPerReq;
  const uint64_t base = (uint64_t)t * kScanTile + threadIdx.x * 16u;
  uint32_t x[16], sum = 0;
  for (int k = 0; k < 16; ++k) {
    x[k] = base + k < n ? in[base + k] : 0u;
    sum += x[k];
  }
  uint32_t agg;
  const uint32_t pre = block_excl_u32(sum, agg);
  __shared__ uint32_t s_prefix;
  __shared__ int s_bad;
  if (threadIdx.x == 0) {
    uint32_t prefix = 0;
    int bad = 0;
    if (t == 0) {
      __hip_atomic_store(lb, (2ull << 32) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(lb + t, (1ull << 32) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t i = t - 1, polls = 0;
      for (;;) {
        uint64_t w = __hip_atomic_load(lb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (((w >> 32) & 3u) == 0) {
          if (++polls > p.spin_limit) { bad = 1; break; }
          __builtin_amdgcn_s_sleep(1);
          w = __hip_atomic_load(lb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (bad) break;
        prefix += (uint32_t)w;
        if (((w >> 32) & 3u) == 2u) break;
        --i;
      }
      if (!bad)
        __hip_atomic_store(lb + t, (2ull << 32) | (prefix + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_prefix = prefix;
    s_bad = bad;
  }
  __syncthreads();
  if (s_bad) return false;
  uint32_t run = s_prefix + pre;
  for (int k = 0; k < 16; ++k) {
    run += x[k];
    if (base + k < n) out[base + k] = run;
  }
  return true;
}

// workgroup-level BFS for small graphs (level-synchronous, queue in scratch)
__device__ void tile_bfs(Slot &s) {
  const uint32_t n = (uint32_t)s.r.n, src = s.r.arg;
  const uint32_t *row = (const uint32_t *)s.r.in, *col = (const uint32_t *)s.r.aux0;
  int32_t *level = (int32_t *)s.r.out;
  uint32_t *q = (uint32_t *)s.r.aux1;
  __shared__ uint32_t s_cnt[2];
  for (uint32_t v = threadIdx.x; v < n; v += kWg) level[v] = v == src ? 0 : kInf;
  if (threadIdx.x == 0) { q[0] = src; s_cnt[0] = 1; s_cnt[1] = 0; }
  __syncthreads();
  for (uint32_t k = 0;; ++k) {
    const uint32_t nin = s_cnt[k & 1];
    if (nin == 0) break;
    const uint32_t *qin = q + (size_t)(k & 1) * n;
    uint32_t *qout = q + (size_t)((k + 1) & 1) * n;
    for (uint32_t i = threadIdx.x; i < nin; i += kWg) {
      const uint32_t u = qin[i];
      for (uint32_t e = row[u]; e < row[u + 1]; ++e) {
        const uint32_t v = col[e];
        int32_t expected = kInf;
        if (level[v] == kInf &&
            __hip_atomic_compare_exchange_strong(level + v, &expected, (int32_t)k + 1, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
          qout[atomicAdd(&s_cnt[(k + 1) & 1], 1u)] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_cnt[k & 1] = 0;
    __syncthreads();
  }
}

// workgroup-level PageRank for small graphs: normalized pull iterations (arg
// of them), FP32, fixed summation order
__device__ void tile_pr(Slot &s) {
  const uint32_t n = (uint32_t)s.r.n, iters = s.r.arg;
  const uint32_t *rrow = (const uint32_t *)s.r.in, *rcol = (const uint32_t *)s.r.aux0;
  const uint32_t *row = (const uint32_t *)s.r.aux2;
  float *x = (float *)s.r.out;
  float *contrib = (float *)s.r.aux1, *y = contrib + n;
  const float a = 0.85f;
  __shared__ float s_dang[kWg / 64];
  for (uint32_t v = threadIdx.x; v < n; v += kWg) x[v] = 1.0f / (float)n;
  __syncthreads();
  for (uint32_t it = 0; it < iters; ++it) {
    float dl = 0.f;
    for (uint32_t v = threadIdx.x; v < n; v += kWg) {
      const uint32_t d = row[v + 1] - row[v];
      contrib[v] = d ? x[v] / (float)d : 0.f;
      if (!d) dl += x[v];
    }
    for (int k = 32; k >= 1; k >>= 1) dl += __shfl_xor(dl, k);
    if (lane_id() == 0) s_dang[threadIdx.x >> 6] = dl;
    __syncthreads();
    float D = 0.f;
    for (int w = 0; w < kWg / 64; ++w) D += s_dang[w];
    const float base = (1.f - a) / (float)n + a * D / (float)n;
    for (uint32_t v = threadIdx.x; v < n; v += kWg) {
      float acc = 0.f;
      for (uint32_t e = rrow[v]; e < rrow[v + 1]; ++e) acc += contrib[rcol[e]];
      y[v] = base + a * acc;
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < n; v += kWg) x[v] = y[v];
    __syncthreads();
  }
}

__device__ uint32_t ntiles_of(const Req &r) {
  switch (r.type) {
    case kReqReduce: return r.n ? (uint32_t)((r.n + kRedTile - 1) / kRedTile) : 1u;
    case kReqScan: return r.n ? (uint32_t)((r.n + kScanTile - 1) / kScanTile) : 1u;
    default: return 1;
  }
}

__global__ void __launch_bounds__(kWg) k_serve(ServeArgs p) {
  __shared__ uint64_t s_task;  // 0 none; 1<<62 | slot | tile<<32 : tile; 2<<62 | index : admission
  __shared__ int s_exit, s_ok;
  uint32_t idle = 0, iter = 0;
  const bool poller = blockIdx.x < kPollers;
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
      const uint32_t head = s_task ? 0u : ld_relaxed(&p.st->qhead);
      if (!s_task && head < ld_relaxed(&p.st->qtail)) {
        uint32_t h = head;
        if (__hip_atomic_compare_exchange_strong(&p.st->qhead, &h, head + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          uint64_t *ent = p.tiles + (head % kQcap);
          uint64_t e;
          uint32_t polls = 0;
          while ((e = __hip_atomic_load(ent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
            if (++polls > p.spin_limit) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (e) {
            __hip_atomic_store(ent, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            fence_acquire();
            s_task = (1ull << 62) | (e - 1);
          } else {
            s_exit = 2;
          }
        }
      }
      if (!s_task && !s_exit && ld_relaxed(&p.st->stop) && ld_relaxed(&p.st->qhead) >= ld_relaxed(&p.st->qtail))
        s_exit = 1;
      if (poller && (iter & 63u) == 0 && ld_sys((const uint32_t *)p.abort)) st_relaxed(&p.st->aborted, 1u);
      if (ld_relaxed(&p.st->aborted)) s_exit = 3;
    }
    __syncthreads();
    const uint64_t task = s_task;
    const int ex = s_exit;
    __syncthreads();
    if (ex) return;
    if (!task) {
      if (++idle > p.spin_limit) return;
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    idle = 0;
    if ((task >> 62) == 2) {
      // admission of request index c into slot c % cap
      const uint32_t c = (uint32_t)task;
      const uint32_t slot = c % p.cap;
      Slot &s = p.slots[slot];
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope: the host's descriptor
        const Req *hr = p.ring + slot;
        s.r = *hr;
        s.index = c;
        if (s.r.type == kReqStop) {
          st_relaxed(&p.st->stop, 1u);
          __hip_atomic_store(&p.done[slot].done, c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
          s_ok = 0;
        } else if (!budget_acquire(p, s.r.mem_units)) {
          st_relaxed(&p.st->stop, 1u);
          s_ok = -1;
        } else {
          p.done[slot].t_admit = wall_clock64();
          s.ntiles = ntiles_of(s.r);
          s.remaining = s.ntiles;
          s.acc = 0;
          s_ok = 1;
          // room in the tile queue (bounded wait)
          uint32_t polls = 0;
          while (ld_relaxed(&p.st->qtail) - ld_relaxed(&p.st->qhead) + s.ntiles > kQcap) {
            if (++polls > p.spin_limit) { s_ok = -1; break; }
            __builtin_amdgcn_s_sleep(4);
          }
        }
      }
      __syncthreads();
      if (s_ok < 0) return;
      if (s_ok == 0) continue;
      if (s.r.type == kReqScan) {
        uint64_t *lb = p.lookback + (size_t)slot * kMaxTilesPerReq;
        for (uint32_t i = threadIdx.x; i < s.ntiles; i += kWg) lb[i] = 0;
      }
      wait_stores();
      __syncthreads();
      if (threadIdx.x == 0) {
        fence_release();
        const uint32_t at = add_relaxed(&p.st->qtail, s.ntiles);
        for (uint32_t i = 0; i < s.ntiles; ++i)
          __hip_atomic_store(p.tiles + ((at + i) % kQcap), ((uint64_t)slot | ((uint64_t)i << 32)) + 1,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      continue;
    }
    // tile execution
    const uint64_t payload = task & ((1ull << 62) - 1);
    const uint32_t slot = (uint32_t)payload, t = (uint32_t)(payload >> 32);
    Slot &s = p.slots[slot];
    const uint32_t type = s.r.type;
    bool ok = true;
    if (type == kReqReduce) tile_reduce(s, t);
    else if (type == kReqScan) ok = tile_scan(p, s, slot, t);
    else if (type == kReqBfs) tile_bfs(s);
    else if (type == kReqPr) tile_pr(s);
    wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
      fence_release();
      if (add_acq_rel(&s.remaining, ~0u) == 1u) {  // last tile of the request
        if (type == kReqReduce) ((int32_t *)s.r.out)[0] = (int32_t)ld_relaxed(&s.acc);
        wait_stores();
        fence_release();
        Done *d = p.done + slot;
        d->status = ok ? 0u : 5u;
        d->t_end = wall_clock64();
        budget_release(p, s.r.mem_units);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __hip_atomic_store(&d->done, s.index + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    __syncthreads();
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
  if (!out || ring_capacity == 0 || ring_capacity > 4096 || budget_units == 0) return HS_ERR_INVALID;
  const void *k = (const void *)k_serve;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 1);
  if (hs_host::check_residency(k, num_wgs, kWg, 0) != HS_OK) return HS_ERR_RESIDENCY;
  Server *s = new Server();
  // a stream of its own that does not synchronize with the null stream, so
  // the host can copy results while the loop runs (default priority)
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { delete s; return HS_ERR_HIP; }
  if (hipHostMalloc((void **)&s->ring, sizeof(Req) * ring_capacity, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&s->done, sizeof(Done) * ring_capacity, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&s->abort, 256, hipHostMallocCoherent) != hipSuccess) {
    delete s;
    return HS_ERR_NOMEM;
  }
  memset(s->ring, 0, sizeof(Req) * ring_capacity);
  memset(s->done, 0, sizeof(Done) * ring_capacity);
  memset(s->abort, 0, 256);
  const size_t head = sizeof(ServeState) + sizeof(Slot) * ring_capacity + sizeof(uint64_t) * kQcap;
  const size_t bytes = head + sizeof(uint64_t) * (size_t)ring_capacity * kMaxTilesPerReq;
  if (hipMalloc(&s->dev, bytes) != hipSuccess) { delete s; return HS_ERR_NOMEM; }
  std::vector<char> zero(head, 0);
  ServeState init{};
  init.avail = budget_units;
  memcpy(zero.data(), &init, sizeof(init));
  HS_TRY_HIP(hipMemcpyAsync(s->dev, zero.data(), head, hipMemcpyHostToDevice, s->stream));
  HS_TRY_HIP(hipStreamSynchronize(s->stream));
  char *w = (char *)s->dev;
  ServeArgs &a = s->args;
  a.st = (ServeState *)w; w += sizeof(ServeState);
  a.slots = (Slot *)w; w += sizeof(Slot) * ring_capacity;
  a.tiles = (uint64_t *)w; w += sizeof(uint64_t) * kQcap;
  a.lookback = (uint64_t *)w;
  a.ring = s->ring;
  a.done = s->done;
  a.abort = s->abort;
  a.cap = ring_capacity;
  a.budget = budget_units;
  a.spin_limit = 1u << 28;
  void *args[] = {(void *)&s->args};
  HS_TRY_HIP(hipLaunchKernel(k, dim3(num_wgs), dim3(kWg), args, 0, s->stream));
  *out = (hs_server *)s;
  return HS_OK;
}

int hs_serve_submit(hs_server *srv, uint32_t type, uint32_t mem_units, uint32_t arg, uint64_t in, uint64_t out,
                    uint64_t n, uint64_t aux0, uint64_t aux1, uint64_t aux2, uint32_t *index_out) {
  Server *s = (Server *)srv;
  if (!s || !index_out) return HS_ERR_INVALID;
  if (type != kReqReduce && type != kReqScan && type != kReqBfs && type != kReqPr && type != kReqStop)
    return HS_ERR_INVALID;
  if (type == kReqScan && (n + kScanTile - 1) / kScanTile > kMaxTilesPerReq) return HS_ERR_INVALID;
  if ((type == kReqBfs || type == kReqPr) && (n == 0 || n > 0xFFFFFFFFull)) return HS_ERR_INVALID;
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
  __atomic_store_n(&s->done[slot].done, 0u, __ATOMIC_RELAXED);
  __atomic_store_n(&r->seq, idx + 1, __ATOMIC_RELEASE);
  s->next = idx + 1;
  *index_out = idx;
  return HS_OK;
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
