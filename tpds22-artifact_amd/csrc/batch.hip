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
#include "hs_wave.hpp"

using namespace hs;

namespace {

constexpr int kWg = 256;
constexpr int kWaves = kWg / 64;
// Tiles of 32768 elements (128 KiB): each wave holds 8192 elements in 128
// registers per lane (2 waves per SIMD).  Measured on 2^28 elements: 8192-element
// tiles 0.57 ms, 16384 0.48 ms, 32768 0.43 ms (the tile counter is one address,
// each atomic on it serializes for ~11 ns, and each tile pays one look-back).
constexpr int kSegs = 32;                  // 256-element segments per wave and tile
constexpr int kWaveSpan = 256 * kSegs;     // 8192 elements per wave
constexpr int kTile = kWaves * kWaveSpan;  // 32768 elements per tile
constexpr uint64_t kFlagAgg = 1, kFlagInc = 2;
constexpr uint32_t kSpinLimit = 1u << 26;

__device__ __forceinline__ uint64_t pack(uint64_t flag, uint32_t v) { return (flag << 32) | v; }

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

// Tiles of a segment [sb, se) are laid on the grid aligned down to 16 bytes:
// tile j covers [A + j*kTile, A + (j+1)*kTile) with A = sb & ~3, clipped to the
// segment, so interior tiles use 16-byte loads and stores whatever the
// segment's alignment.  Inside a tile, wave v owns [v*kWaveSpan, +kWaveSpan)
// and, in its segment q, lane l holds the 4 elements at q*256 + 4l: every
// instruction of a wave touches 1 KiB of contiguous memory.
__host__ __device__ __forceinline__ uint64_t seg_tiles(uint64_t sb, uint64_t se) {
  return se > sb ? (se - (sb & ~3ull) + kTile - 1) / kTile : 0;
}

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

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// this wave's kWaveSpan elements of a tile starting at `base`: x[q][0..3] of lane l
// are elements base + q*256 + 4l + 0..3; elements outside [lo, hi) read as 0.
// Every element is read once and written once, so interior tiles use
// non-temporal (streaming) 16-byte loads and stores: on a 2^28-element scan
// 0.412 ms against 0.433 ms with plain ones.
__device__ __forceinline__ void wave_load(const uint32_t *in, uint64_t base, uint64_t lo, uint64_t hi,
                                          uint32_t lane, uint32_t (&x)[kSegs][4]) {
  if (base >= lo && base + kWaveSpan <= hi) {
    const u32x4 *p = reinterpret_cast<const u32x4 *>(in + base) + lane;
#pragma unroll
    for (int q = 0; q < kSegs; ++q) {
      const u32x4 v = __builtin_nontemporal_load(p + q * 64);
      x[q][0] = v.x; x[q][1] = v.y; x[q][2] = v.z; x[q][3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < kSegs; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t i = base + (uint64_t)q * 256 + 4u * lane + j;
        x[q][j] = (i >= lo && i < hi) ? in[i] : 0u;
      }
  }
}

__device__ __forceinline__ void wave_store(uint32_t *out, uint64_t base, uint64_t lo, uint64_t hi, uint32_t lane,
                                           const uint32_t (&y)[kSegs][4]) {
  if (base >= lo && base + kWaveSpan <= hi) {
    u32x4 *o = reinterpret_cast<u32x4 *>(out + base) + lane;
#pragma unroll
    for (int q = 0; q < kSegs; ++q) {
      const u32x4 v = {y[q][0], y[q][1], y[q][2], y[q][3]};
      __builtin_nontemporal_store(v, o + q * 64);
    }
  } else {
#pragma unroll
    for (int q = 0; q < kSegs; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t i = base + (uint64_t)q * 256 + 4u * lane + j;
        if (i >= lo && i < hi) out[i] = y[q][j];
      }
  }
}

__device__ __forceinline__ uint32_t ld_status_value(const uint64_t *p, uint64_t &flag) {
  const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  flag = (w >> 32) & 3u;
  return (uint32_t)w;
}

// Exclusive prefix of tile t (first tile of its segment: `first`) from the
// status words of its predecessors, 64 at a time: lane i looks at tile t-1-i;
// tiles before `first` count as inclusive zeros.  Waits until every lane up to
// the nearest inclusive word is published.  Called by one whole wave.
__device__ __forceinline__ int lookback(const BatchArgs &p, uint64_t t, uint64_t first, uint32_t lane,
                                        uint32_t &prefix) {
  prefix = 0;
  int64_t i = (int64_t)t - 1;
  uint32_t polls = 0;
  for (;;) {
    const int64_t idx = i - (int64_t)lane;
    uint64_t flag = kFlagInc;
    uint32_t v = 0;
    if (idx >= (int64_t)first) v = ld_status_value(p.status + idx, flag);
    const uint64_t incm = __builtin_amdgcn_ballot_w64(flag == kFlagInc);
    const uint64_t readym = __builtin_amdgcn_ballot_w64(flag != 0);
    const uint32_t k = incm ? (uint32_t)__builtin_ctzll(incm) : 63u;
    const uint64_t need = k == 63 ? ~0ull : ((2ull << k) - 1ull);
    if ((readym & need) != need) {
      // give up uniformly across the wave (the error word is read once)
      if (++polls >= kSpinLimit || __builtin_amdgcn_readfirstlane(ld_relaxed(p.error))) return 1;
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += wave_sum(lane <= k ? v : 0u);
    if (incm) return 0;
    i -= 64;
  }
}

__global__ void __launch_bounds__(kWg) k_scan_batch(BatchArgs p) {
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ uint32_t s_wave[kWaves];
  __shared__ int s_bad;
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) s_tile = add_relaxed(p.counter, 1u);
    __syncthreads();
    const uint64_t t = s_tile;
    __syncthreads();
    if (t >= p.ntiles) return;
    const uint32_t s = seg_of(p, t);
    const uint64_t first = p.tile_start[s];
    const uint64_t j = t - first;  // tile within the segment
    const uint64_t sb = p.offsets[s], se = p.offsets[s + 1];
    const uint64_t base = (sb & ~3ull) + j * kTile + (uint64_t)wv * kWaveSpan;
    uint32_t x[kSegs][4];
    wave_load(p.in, base, sb, se, lane, x);
    // local scan of the wave's kWaveSpan elements (inclusive), wave total
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kSegs; ++q) {
      const uint32_t v4 = x[q][0] + x[q][1] + x[q][2] + x[q][3];
      const uint32_t inc = wave_incl_scan(v4);
      uint32_t pre = run + inc - v4;
      run += __builtin_amdgcn_readlane(inc, 63);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t v = x[q][k];
        if (p.inclusive) { pre += v; x[q][k] = pre; }
        else { x[q][k] = pre; pre += v; }
      }
    }
    if (lane == 0) s_wave[wv] = run;
    __syncthreads();
    uint32_t wave_before = 0, agg = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      wave_before += k < (int)wv ? s_wave[k] : 0u;
      agg += s_wave[k];
    }
    if (wv == 0) {
      uint32_t prefix = 0;
      int bad = 0;
      if (j == 0) {
        if (lane == 0)
          __hip_atomic_store(p.status + t, pack(kFlagInc, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        if (lane == 0)
          __hip_atomic_store(p.status + t, pack(kFlagAgg, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bad = lookback(p, t, first, lane, prefix);
        if (lane == 0) {
          if (bad) st_relaxed(p.error, 1u);
          else
            __hip_atomic_store(p.status + t, pack(kFlagInc, prefix + agg), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      if (lane == 0) { s_prefix = prefix; s_bad = bad; }
    }
    __syncthreads();
    if (s_bad) return;
    const uint32_t add = s_prefix + wave_before;
#pragma unroll
    for (int q = 0; q < kSegs; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) x[q][k] += add;
    wave_store(p.out, base, sb, se, lane, x);
    __syncthreads();  // s_tile, s_wave and s_prefix are rewritten by the next tile
  }
}

// batched reduce: the tile sums of a segment are added into out[s] with
// device-scope atomics (uint32 wrap: order-independent, so bit-exact)
__global__ void __launch_bounds__(kWg) k_reduce_batch(BatchArgs p) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave[kWaves];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) s_tile = add_relaxed(p.counter, 1u);
    __syncthreads();
    const uint64_t t = s_tile;
    __syncthreads();
    if (t >= p.ntiles) return;
    const uint32_t s = seg_of(p, t);
    const uint64_t j = t - p.tile_start[s];
    const uint64_t sb = p.offsets[s], se = p.offsets[s + 1];
    const uint64_t base = (sb & ~3ull) + j * kTile + (uint64_t)wv * kWaveSpan;
    uint32_t x[kSegs][4];
    wave_load(p.in, base, sb, se, lane, x);
    uint32_t acc = 0;
#pragma unroll
    for (int q = 0; q < kSegs; ++q) acc += x[q][0] + x[q][1] + x[q][2] + x[q][3];
    acc = wave_sum(acc);
    if (lane == 0) s_wave[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t total = 0;
#pragma unroll
      for (int k = 0; k < kWaves; ++k) total += s_wave[k];
      add_relaxed(p.out + s, total);
    }
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
      acc += seg_tiles(offsets[s], offsets[s + 1]);
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
    ntiles += seg_tiles(h_offsets[i], h_offsets[i + 1]);
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
  // 4 workgroups per CU: a workgroup has one tile in flight, so the streaming
  // rate grows with resident workgroups (measured: 2.36 / 2.72 / 3.05 TB/s at 1 / 2 / 4 per CU)
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 4);
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
