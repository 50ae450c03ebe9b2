// batch.hip -- many uneven scans / reduces in one persistent launch.
//
// No reference counterpart: the reference launches one scan or reduce per
// binary (scan/main.cu, Reduce/test_reduce.cu).  Here the segments of a batch
// are cut into tiles of kTile elements; workgroups take tiles in increasing
// global order from one counter (work stealing, so short and long segments mix
// freely), and the carry of each tile comes from a decoupled look-back over its
// predecessors of the same segment (Merrill & Garland, "Single-pass Parallel
// Prefix Scan with Decoupled Look-back", 2016).  A tile publishes a 64-bit
// status word {flag:2, value:32} with one device-scope store; the value
// travels inside the word, so no cache writeback or invalidate is needed per
// tile.  The status words are cleared by a fill kernel before every launch.
// Every tile is read once and written once: 8 bytes per element for the scan.
// Look-back waits are bounded: a wait that exceeds the spin limit records an
// error word and the workgroup exits.
#include <hip/hip_runtime.h>

#include "hs_internal.hpp"
#include "wave.hpp"

using namespace hs;

namespace {

constexpr int kWg = 256;
constexpr int kItems = 16;
constexpr int kTile = kWg * kItems;  // 4096 elements
constexpr uint64_t kFlagAgg = 1, kFlagInc = 2;
constexpr uint32_t kSpinLimit = 1u << 26;

__device__ __forceinline__ uint64_t pack(uint64_t flag, uint32_t v) { return (flag << 32) | v; }

// exclusive block scan of one value per thread + block total
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t &total) {
  __shared__ uint32_t ws[kWg / 64];
  const uint32_t inc = wave_incl_scan(v);
  if (lane_id() == 63) ws[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint32_t before = 0, t = 0;
#pragma unroll
  for (int i = 0; i < kWg / 64; ++i) {
    before += i < (int)(threadIdx.x >> 6) ? ws[i] : 0u;
    t += ws[i];
  }
  __syncthreads();
  total = t;
  return before + inc - v;
}

struct BatchArgs {
  const uint32_t *in;
  uint32_t *out;               // scan output (same layout as in) or per-segment sums (reduce)
  const uint64_t *offsets;     // [nseg+1] element offsets of the segments
  const uint64_t *tile_start;  // [nseg+1] first global tile of each segment
  uint32_t nseg;
  uint64_t ntiles;
  uint64_t *status;            // [ntiles] look-back words
  uint32_t *counter;           // tile counter (zeroed by the prep kernel)
  uint32_t *error;             // set when a look-back wait gives up
  int inclusive;
};

// segment of global tile t: binary search in tile_start
__device__ __forceinline__ uint32_t seg_of(const BatchArgs &p, uint64_t t) {
  uint32_t lo = 0, hi = p.nseg;  // tile_start[lo] <= t < tile_start[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (p.tile_start[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kWg) k_scan_batch(BatchArgs p) {
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ int s_bad;
  for (;;) {
    if (threadIdx.x == 0) s_tile = add_relaxed(p.counter, 1u);
    __syncthreads();
    const uint64_t t = s_tile;
    __syncthreads();
    if (t >= p.ntiles) return;
    const uint32_t s = seg_of(p, t);
    const uint64_t j = t - p.tile_start[s];  // tile within the segment
    const uint64_t sb = p.offsets[s], se = p.offsets[s + 1];
    const uint64_t base = sb + j * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t x[kItems];
    const bool full = base + kItems <= se && ((sb & 3u) == 0);
    if (full) {
      const uint4 *p4 = reinterpret_cast<const uint4 *>(p.in + base);
#pragma unroll
      for (int k = 0; k < kItems / 4; ++k) {
        const uint4 v = p4[k];
        x[4 * k] = v.x; x[4 * k + 1] = v.y; x[4 * k + 2] = v.z; x[4 * k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kItems; ++k) x[k] = base + k < se ? p.in[base + k] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) tsum += x[k];
    uint32_t agg;
    const uint32_t pre = block_excl(tsum, agg);
    if (threadIdx.x == 0) {
      uint32_t prefix = 0;
      int bad = 0;
      if (j == 0) {
        __hip_atomic_store(p.status + t, pack(kFlagInc, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(p.status + t, pack(kFlagAgg, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // look back over the predecessors of this segment; every one of them
        // was handed to a running workgroup before this tile, so each wait ends
        uint64_t i = t - 1;
        uint32_t polls = 0;
        for (;;) {
          uint64_t w = __hip_atomic_load(p.status + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          while (((w >> 32) & 3u) == 0) {
            if (++polls >= kSpinLimit || ld_relaxed(p.error)) { bad = 1; break; }
            __builtin_amdgcn_s_sleep(1);
            w = __hip_atomic_load(p.status + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          if (bad) break;
          prefix += (uint32_t)w;
          if (((w >> 32) & 3u) == kFlagInc) break;
          --i;
        }
        if (bad) st_relaxed(p.error, 1u);
        else
          __hip_atomic_store(p.status + t, pack(kFlagInc, prefix + agg), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
      s_prefix = prefix;
      s_bad = bad;
    }
    __syncthreads();
    if (s_bad) return;
    uint32_t run = s_prefix + pre;
    uint32_t y[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      if (p.inclusive) { run += x[k]; y[k] = run; }
      else { y[k] = run; run += x[k]; }
    }
    if (full) {
      uint4 *o4 = reinterpret_cast<uint4 *>(p.out + base);
#pragma unroll
      for (int k = 0; k < kItems / 4; ++k) o4[k] = make_uint4(y[4 * k], y[4 * k + 1], y[4 * k + 2], y[4 * k + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < kItems; ++k)
        if (base + k < se) p.out[base + k] = y[k];
    }
  }
}

// batched reduce: the tile sums of a segment are added into out[s] with
// device-scope atomics (uint32 wrap: order-independent, so bit-exact)
__global__ void __launch_bounds__(kWg) k_reduce_batch(BatchArgs p) {
  __shared__ uint32_t s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = add_relaxed(p.counter, 1u);
    __syncthreads();
    const uint64_t t = s_tile;
    __syncthreads();
    if (t >= p.ntiles) return;
    const uint32_t s = seg_of(p, t);
    const uint64_t j = t - p.tile_start[s];
    const uint64_t sb = p.offsets[s], se = p.offsets[s + 1];
    const uint64_t base = sb + j * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t acc = 0;
    if (base + kItems <= se && (sb & 3u) == 0) {
      const uint4 *p4 = reinterpret_cast<const uint4 *>(p.in + base);
#pragma unroll
      for (int k = 0; k < kItems / 4; ++k) {
        const uint4 v = p4[k];
        acc += v.x + v.y + v.z + v.w;
      }
    } else {
      for (int k = 0; k < kItems; ++k)
        if (base + k < se) acc += p.in[base + k];
    }
    uint32_t total;
    block_excl(acc, total);
    if (threadIdx.x == 0) add_relaxed(p.out + s, total);
    // Keep this barrier: without it the thread-0 region above and the
    // thread-0 tile claim at the loop head are merged across the back edge,
    // and the AMDGPU structurizer then runs the other lanes around the loop
    // (re-reading a stale tile index) before thread 0 claims the next tile,
    // which never ends.
    __syncthreads();
  }
}

// tile starts (serial prefix of per-segment tile counts, nseg <= 65536) and
// reset of the counter, the error word and (reduce) the sums
__global__ void k_batch_prep(const uint64_t *offsets, uint32_t nseg, uint64_t *tile_start, uint32_t *counter,
                             uint32_t *error, uint32_t *sums, int zero_sums) {
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
      tile_start[s] = acc;
      const uint64_t len = offsets[s + 1] - offsets[s];
      acc += (len + kTile - 1) / kTile;
    }
    tile_start[nseg] = acc;
    *counter = 0;
    *error = 0;
  }
  if (zero_sums)
    for (uint32_t s = threadIdx.x; s < nseg; s += blockDim.x) sums[s] = 0;
}

__global__ void k_clear_u64(uint64_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

uint64_t count_tiles(const uint64_t *h_offsets, uint32_t nseg, bool *ok) {
  uint64_t ntiles = 0;
  *ok = true;
  for (uint32_t i = 0; i < nseg; ++i) {
    if (h_offsets[i + 1] < h_offsets[i]) { *ok = false; return 0; }
    ntiles += (h_offsets[i + 1] - h_offsets[i] + kTile - 1) / kTile;
  }
  return ntiles;
}

size_t tile_start_bytes(uint32_t nseg) { return ((sizeof(uint64_t) * (nseg + 1) + 255) / 256) * 256; }

int launch_batch(bool scan, const uint32_t *d_in, uint32_t *d_out, const uint64_t *d_offsets,
                 const uint64_t *h_offsets, uint32_t nseg, int inclusive, void *d_work, int num_wgs,
                 hipStream_t s) {
  if (!d_in || !d_out || !d_offsets || !h_offsets || !d_work || nseg == 0 || nseg > 65536) return HS_ERR_INVALID;
  bool ok;
  const uint64_t ntiles = count_tiles(h_offsets, nseg, &ok);
  if (!ok || ntiles >= 0xFFFFFFFFull) return HS_ERR_INVALID;
  const void *k = scan ? (const void *)k_scan_batch : (const void *)k_reduce_batch;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 2);
  if (hs_host::check_residency(k, num_wgs, kWg, 0) != HS_OK) return HS_ERR_RESIDENCY;
  char *w = (char *)d_work;
  BatchArgs p{};
  p.in = d_in; p.out = d_out; p.offsets = d_offsets; p.nseg = nseg; p.ntiles = ntiles;
  p.inclusive = inclusive;
  p.counter = (uint32_t *)w;
  p.error = (uint32_t *)(w + 128);
  w += 256;
  p.tile_start = (const uint64_t *)w;
  w += tile_start_bytes(nseg);
  p.status = (uint64_t *)w;
  hipLaunchKernelGGL(k_batch_prep, dim3(1), dim3(256), 0, s, d_offsets, nseg, (uint64_t *)p.tile_start,
                     p.counter, p.error, d_out, scan ? 0 : 1);
  if (scan && ntiles) {
    const int blocks = (int)((ntiles + 255) / 256 > 1024 ? 1024 : (ntiles + 255) / 256);
    hipLaunchKernelGGL(k_clear_u64, dim3(blocks), dim3(256), 0, s, p.status, ntiles);
  }
  HS_TRY_HIP(hipGetLastError());
  if (ntiles == 0) return HS_OK;
  void *args[] = {(void *)&p};
  HS_TRY_HIP(hipLaunchKernel(k, dim3(num_wgs), dim3(kWg), args, 0, s));
  return HS_OK;
}

}  // namespace

extern "C" {

int hs_batch_work_size(const uint64_t *h_offsets, uint32_t nseg, uint64_t *bytes_out) {
  if (!h_offsets || !bytes_out || nseg == 0 || nseg > 65536) return HS_ERR_INVALID;
  bool ok;
  const uint64_t ntiles = count_tiles(h_offsets, nseg, &ok);
  if (!ok) return HS_ERR_INVALID;
  *bytes_out = 256 + tile_start_bytes(nseg) + sizeof(uint64_t) * (ntiles + 1);
  return HS_OK;
}

// status of the last batched launch that used d_work: HS_OK or HS_ERR_TIMEOUT
int hs_batch_status(const void *d_work, void *stream) {
  if (!d_work) return HS_ERR_INVALID;
  uint32_t err = 0;
  HS_TRY_HIP(hipMemcpyAsync(&err, (const char *)d_work + 128, sizeof(err), hipMemcpyDeviceToHost,
                            (hipStream_t)stream));
  HS_TRY_HIP(hipStreamSynchronize((hipStream_t)stream));
  return err ? HS_ERR_TIMEOUT : HS_OK;
}

int hs_scan_batch_u32(const uint32_t *d_in, uint32_t *d_out, const uint64_t *d_offsets, const uint64_t *h_offsets,
                      uint32_t nseg, int inclusive, void *d_work, int num_wgs, void *stream) {
  if (((uintptr_t)d_in & 15u) || ((uintptr_t)d_out & 15u)) return HS_ERR_INVALID;
  return launch_batch(true, d_in, d_out, d_offsets, h_offsets, nseg, inclusive, d_work, num_wgs,
                      (hipStream_t)stream);
}

int hs_reduce_batch_i32(const int32_t *d_in, int32_t *d_sums, const uint64_t *d_offsets, const uint64_t *h_offsets,
                        uint32_t nseg, void *d_work, int num_wgs, void *stream) {
  if ((uintptr_t)d_in & 15u) return HS_ERR_INVALID;
  return launch_batch(false, (const uint32_t *)d_in, (uint32_t *)d_sums, d_offsets, h_offsets, nseg, 0, d_work,
                      num_wgs, (hipStream_t)stream);
}

}  // extern "C"
