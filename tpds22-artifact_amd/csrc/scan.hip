// scan.hip -- persistent multi-phase u32 prefix sum with grid barriers.
//
// Replaces gpu_prescan / gpu_add_block_sums / sum_scan_blelloch
// (scan/scan_gsrb.cu:440-804).  The reference runs a padded-LDS Blelloch scan of
// 64 elements per 32-thread block, a barrier per level, and two more host
// launches, and is correct only up to 64^3 elements.  Here one co-resident grid
// processes the input in rounds sized to stay resident in the Infinity Cache:
//   phase 1  each workgroup sums its contiguous chunk of the round (one HBM read)
//   barrier  (any hs_barrier_kind)
//   phase 2  each workgroup folds the partials before it (its carry) and the
//            round total
//   phase 3  each workgroup re-reads its chunk (Infinity-Cache hit), scans it
//            tile by tile (16 elements per thread, DPP wave scan, LDS block scan)
//            and writes the result with the carry.
// Partials are double-buffered by round parity, so one barrier per round
// suffices.  uint32 wrap arithmetic: bit-exact for any grid and round size.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hs_internal.hpp"

using namespace hs;

// Partials of hs_scan_u32_noalloc: 2 words per workgroup (double-buffered by
// round parity), for grids of up to kNoAllocMaxWgs workgroups (32 KiB).  Each
// device holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocMaxWgs = 4096;
static __device__ __align__(256) uint32_t scratch[2 * kNoAllocMaxWgs];

namespace {

constexpr int kWg = 256;
constexpr int kItems = 16;                 // elements per thread per tile
constexpr int kTile = kWg * kItems;        // 4096 elements

// inclusive wave64 scan (DPP row_shr + row_bcast)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);
  return v;
}

// exclusive block scan of one value per thread; returns the exclusive prefix and
// the block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t wsum[kWg / 64];
  const uint32_t inc = wave_incl_scan(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWg / 64; ++i) {
    const uint32_t s = wsum[i];
    before += i < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
  uint32_t total;
  block_excl_scan(v, total);
  return total;
}

// Each workgroup's chunk of a round is split into 4 contiguous wave
// sub-chunks, so phase 3 needs no workgroup barrier per tile: a wave scans its
// own sub-chunk 1024 elements (a wave-tile) at a time, its carry being the
// chunk carry plus the sums of the waves before it (from phase 1).
// Wave-tile layout: 4 segments of 256 elements; in segment s lane l holds
// elements s*256 + 4l .. 4l+3, so every load and store instruction of a wave
// covers 1 KiB of contiguous memory.  Each segment is one DPP wave scan.
constexpr int kSeg = 4;

__device__ __forceinline__ void wt_load(const uint32_t *in, uint64_t tile, uint32_t lane,
                                        uint32_t (&x)[kSeg][4]) {
  const uint4 *p = reinterpret_cast<const uint4 *>(in + tile) + lane;
#pragma unroll
  for (int s = 0; s < kSeg; ++s) {
    const uint4 v = p[s * 64];
    x[s][0] = v.x; x[s][1] = v.y; x[s][2] = v.z; x[s][3] = v.w;
  }
}

// elements at or beyond `end` read as 0
__device__ __forceinline__ void wt_load_part(const uint32_t *in, uint64_t tile, uint64_t end, uint32_t lane,
                                             uint32_t (&x)[kSeg][4]) {
#pragma unroll
  for (int s = 0; s < kSeg; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t i = tile + (uint64_t)s * 256 + 4u * lane + j;
      x[s][j] = i < end ? in[i] : 0u;
    }
}

// scans the wave-tile in place (x becomes the output) and advances run
template <bool INCLUSIVE>
__device__ __forceinline__ void wt_scan(uint32_t (&x)[kSeg][4], uint32_t &run) {
#pragma unroll
  for (int s = 0; s < kSeg; ++s) {
    const uint32_t t = x[s][0] + x[s][1] + x[s][2] + x[s][3];
    const uint32_t inc = wave_incl_scan(t);
    uint32_t pre = run + inc - t;
    run += __builtin_amdgcn_readlane(inc, 63);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t v = x[s][j];
      if (INCLUSIVE) { pre += v; x[s][j] = pre; }
      else { x[s][j] = pre; pre += v; }
    }
  }
}

// output stores are non-temporal (streaming: nothing reads them back in the
// launch); the input's second read stays a plain load, meant to hit the
// Infinity Cache the first read filled
__device__ __forceinline__ void wt_store(uint32_t *out, uint64_t tile, uint32_t lane, const uint32_t (&y)[kSeg][4]) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4 *o = reinterpret_cast<u32x4 *>(out + tile) + lane;
#pragma unroll
  for (int s = 0; s < kSeg; ++s) {
    const u32x4 v = {y[s][0], y[s][1], y[s][2], y[s][3]};
    __builtin_nontemporal_store(v, o + s * 64);
  }
}

__device__ __forceinline__ void wt_store_part(uint32_t *out, uint64_t tile, uint64_t end, uint32_t lane,
                                              const uint32_t (&y)[kSeg][4]) {
#pragma unroll
  for (int s = 0; s < kSeg; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t i = tile + (uint64_t)s * 256 + 4u * lane + j;
      if (i < end) out[i] = y[s][j];
    }
}

__device__ __forceinline__ uint32_t wt_sum(const uint32_t (&x)[kSeg][4]) {
  uint32_t t = 0;
#pragma unroll
  for (int s = 0; s < kSeg; ++s) t += x[s][0] + x[s][1] + x[s][2] + x[s][3];
  return t;
}

// per-lane sum of the wave's range [sb, se) (16 x 16 B in flight per lane)
__device__ __forceinline__ uint32_t range_sum(const uint32_t *in, uint64_t sb, uint64_t se, uint32_t lane) {
  uint32_t acc = 0;
  const uint64_t full = sb + (se - sb) / 4 * 4;
  const uint4 *p4 = reinterpret_cast<const uint4 *>(in + sb);
  const uint64_t n4 = (full - sb) / 4;
  uint64_t i = lane;
  for (; i + 15 * 64 < n4; i += 16 * 64) {
    uint4 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p4[i + k * 64];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += v[k].x + v[k].y + v[k].z + v[k].w;
  }
  for (; i < n4; i += 64) {
    const uint4 v = p4[i];
    acc += v.x + v.y + v.z + v.w;
  }
  for (uint64_t j = full + lane; j < se; j += 64) acc += in[j];
  return acc;
}

// Rounds are software-pipelined: while a wave scans and writes its sub-chunk
// of round r (re-read from the Infinity Cache), it also reads and sums its
// sub-chunk of round r+1 (from HBM), so reads and writes stream together and
// each round needs one grid barrier.  Partials are double-buffered by round
// parity.
template <int KIND, bool INCLUSIVE>
__global__ void __launch_bounds__(kWg) k_scan_u32(ArenaView a, const uint32_t *__restrict__ in,
                                                 uint32_t *__restrict__ out, uint64_t n,
                                                 uint32_t *partials, uint64_t round_elems,
                                                 uint32_t ngroups, uint32_t nxcd) {
  constexpr int kWaves = kWg / 64;
  constexpr int kWaveTile = 256 * kSeg;  // 1024 elements
  __shared__ uint32_t s_wave[kWaves];
  GridBarrier gb;
  gb.init(a, KIND, ngroups, nxcd);
  const uint32_t W = gridDim.x, w = blockIdx.x;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t chunk = round_elems / W;  // multiple of kTile
  const uint64_t sub = chunk / kWaves;     // multiple of kWaveTile
  const uint64_t off = (uint64_t)w * chunk + wv * sub;  // this wave's offset inside a round
  auto bounds = [&](uint64_t rb, uint64_t &sb, uint64_t &se) {
    const uint64_t re = rb + round_elems < n ? rb + round_elems : n;
    sb = rb + off;
    se = sb + sub;
    if (sb > re) sb = re;
    if (se > re) se = re;
  };
  // publish this workgroup's total of one round (from the per-wave sums)
  auto publish = [&](uint32_t wave_sum, uint32_t *part, uint32_t &wave_before) {
    if (lane == 0) s_wave[wv] = wave_sum;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      before += k < (int)wv ? s_wave[k] : 0u;
      total += s_wave[k];
    }
    wave_before = before;
    if (threadIdx.x == 0) part[w] = total;
    __syncthreads();  // s_wave is rewritten by the next publish
  };
  uint32_t carry = 0;  // sum of all previous rounds
  uint64_t sb, se;
  bounds(0, sb, se);
  uint32_t wave_before;
  publish(__builtin_amdgcn_readlane(wave_incl_scan(range_sum(in, sb, se, lane)), 63), partials, wave_before);
  uint32_t r = 0;
  for (uint64_t rb = 0; rb < n; rb += round_elems, ++r) {
    if (gb.sync<KIND>()) return;
    const uint32_t *part = partials + (r & 1u) * W;
    uint32_t before = 0, all = 0;
    for (uint32_t j = threadIdx.x; j < W; j += kWg) {
      const uint32_t v = part[j];
      all += v;
      before += j < w ? v : 0u;
    }
    before = block_sum(before);
    all = block_sum(all);
    uint32_t run = carry + before + wave_before;
    const bool next = rb + round_elems < n;
    uint64_t nb = 0, ne = 0;  // this wave's range in round r+1
    if (next) bounds(rb + round_elems, nb, ne);
    uint32_t nacc = 0;        // per-lane sum of [nb, ne)
    const uint64_t nf = (se - sb) / kWaveTile;  // full tiles of round r
    const uint64_t nfn = (ne - nb) / kWaveTile;  // full tiles of round r+1 (<= nf)
    uint64_t t = 0;
    // both rounds' tiles full: one tile of each in flight ahead
    if (nfn > 0) {
      uint32_t x[kSeg][4], y[kSeg][4];
      wt_load(in, sb, lane, x);
      wt_load(in, nb, lane, y);
      for (; t + 1 < nfn; ++t) {
        uint32_t x2[kSeg][4], y2[kSeg][4];
        wt_load(in, sb + (t + 1) * kWaveTile, lane, x2);
        wt_load(in, nb + (t + 1) * kWaveTile, lane, y2);
        nacc += wt_sum(y);
        wt_scan<INCLUSIVE>(x, run);
        wt_store(out, sb + t * kWaveTile, lane, x);
#pragma unroll
        for (int q = 0; q < kSeg; ++q)
#pragma unroll
          for (int j = 0; j < 4; ++j) { x[q][j] = x2[q][j]; y[q][j] = y2[q][j]; }
      }
      nacc += wt_sum(y);
      wt_scan<INCLUSIVE>(x, run);
      wt_store(out, sb + t * kWaveTile, lane, x);
      ++t;
    }
    // remaining full tiles of round r, one ahead
    if (t < nf) {
      uint32_t x[kSeg][4];
      wt_load(in, sb + t * kWaveTile, lane, x);
      for (; t + 1 < nf; ++t) {
        uint32_t x2[kSeg][4];
        wt_load(in, sb + (t + 1) * kWaveTile, lane, x2);
        wt_scan<INCLUSIVE>(x, run);
        wt_store(out, sb + t * kWaveTile, lane, x);
#pragma unroll
        for (int q = 0; q < kSeg; ++q)
#pragma unroll
          for (int j = 0; j < 4; ++j) x[q][j] = x2[q][j];
      }
      wt_scan<INCLUSIVE>(x, run);
      wt_store(out, sb + t * kWaveTile, lane, x);
      ++t;
    }
    // partial last tile of round r (only in the last round)
    if (sb + nf * kWaveTile < se) {
      uint32_t x[kSeg][4];
      wt_load_part(in, sb + nf * kWaveTile, se, lane, x);
      wt_scan<INCLUSIVE>(x, run);
      wt_store_part(out, sb + nf * kWaveTile, se, lane, x);
    }
    if (!next) break;
    // the rest of round r+1 (beyond its full tiles overlapped above)
    nacc += range_sum(in, nb + nfn * kWaveTile, ne, lane);
    carry += all;
    sb = nb;
    se = ne;
    publish(__builtin_amdgcn_readlane(wave_incl_scan(nacc), 63), partials + ((r + 1) & 1u) * W, wave_before);
  }
}

template <bool INC>
const void *scan_kernel(int kind) {
  switch (kind) {
    case HS_BAR_FLAT: return (const void *)k_scan_u32<HS_BAR_FLAT, INC>;
    case HS_BAR_TWO_LEVEL: return (const void *)k_scan_u32<HS_BAR_TWO_LEVEL, INC>;
    case HS_BAR_XCD: return (const void *)k_scan_u32<HS_BAR_XCD, INC>;
    case HS_BAR_XCD_FLAT_RELEASE: return (const void *)k_scan_u32<HS_BAR_XCD_FLAT_RELEASE, INC>;
    case HS_BAR_TREE: return (const void *)k_scan_u32<HS_BAR_TREE, INC>;
    case HS_BAR_LF: return (const void *)k_scan_u32<HS_BAR_LF, INC>;
    case HS_BAR_COOP: return (const void *)k_scan_u32<HS_BAR_COOP, INC>;
    case HS_BAR_NAIVE_ALL: return (const void *)k_scan_u32<HS_BAR_NAIVE_ALL, INC>;
    default: return nullptr;
  }
}

}  // namespace

namespace hs_host {
int prepare_barrier(hs_arena *ar, int kind, hipStream_t s);
int launch_persistent(const void *k, void **args, int kind, int num_wgs, int wg_size, size_t lds,
                      hipStream_t s);
}  // namespace hs_host

// round size: ~64 MiB of input per round by default (two rounds in flight stay
// resident in the 256 MiB Infinity Cache; measured best of 32..1024 MiB),
// rounded to a whole number of tiles per workgroup
static uint64_t scan_round_elems(int num_wgs, uint64_t n, uint64_t target_bytes) {
  const uint64_t per_wg_target = target_bytes / 4 / (uint64_t)num_wgs;
  uint64_t tiles = per_wg_target / kTile;
  if (tiles < 1) tiles = 1;
  // no more tiles per workgroup than the whole input needs
  const uint64_t need = (n + (uint64_t)num_wgs * kTile - 1) / ((uint64_t)num_wgs * kTile);
  if (tiles > need && need > 0) tiles = need;
  return tiles * kTile * (uint64_t)num_wgs;
}

static uint64_t g_scan_round_bytes = 64ull << 20;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device
}  // namespace

namespace hs_host {
// Looks up this device's copy of `scratch` (the device must be current).  Called
// once per device from hs_arena_create, so the no-allocation entry point itself
// makes no runtime call beyond its launches.
int scan_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return HS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return HS_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "libhs: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
            hipGetErrorString(e));
    return HS_ERR_HIP;
  }
  g_scratch_dev[device] = (uint32_t *)p;
  return HS_OK;
}
}  // namespace hs_host

extern "C" {

int hs_scan_scratch_size(int num_wgs, uint64_t n, uint64_t *elems_out) {
  (void)n;
  if (!elems_out) return HS_ERR_INVALID;
  if (num_wgs <= 0) num_wgs = 2 * hs_host::device_cus();
  *elems_out = 2ull * (uint64_t)num_wgs;
  return HS_OK;
}

int hs_scan_set_round_bytes(uint64_t bytes) {
  if (bytes < 4096) return HS_ERR_INVALID;
  g_scan_round_bytes = bytes;
  return HS_OK;
}

int hs_scan_get_round_bytes(uint64_t *bytes_out) {
  if (!bytes_out) return HS_ERR_INVALID;
  *bytes_out = g_scan_round_bytes;
  return HS_OK;
}

int hs_scan_u32(hs_arena *ar, const uint32_t *d_in, uint32_t *d_out, uint64_t n, int inclusive,
                uint32_t *d_scratch, int kind, int num_wgs, void *stream) {
  if (!ar || !d_scratch || (n && (!d_in || !d_out))) return HS_ERR_INVALID;
  if (((uintptr_t)d_in & 15u) || ((uintptr_t)d_out & 15u)) return HS_ERR_INVALID;
  const void *k = inclusive ? scan_kernel<true>(kind) : scan_kernel<false>(kind);
  if (!k) return HS_ERR_INVALID;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 2);
  if (num_wgs > ar->view.max_blocks) return HS_ERR_INVALID;
  int rc = hs_host::check_residency(k, num_wgs, kWg, 0);
  if (rc) return rc;
  if (n == 0) return HS_OK;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = hs_host::prepare_barrier(ar, kind, s))) return rc;
  uint64_t round = scan_round_elems(num_wgs, n, g_scan_round_bytes);
  hs::ArenaView v = ar->view;
  uint32_t ng = (uint32_t)v.num_groups, nx = (uint32_t)hs_host::xcd_groups(num_wgs);
  void *args[] = {(void *)&v, (void *)&d_in, (void *)&d_out, (void *)&n, (void *)&d_scratch,
                  (void *)&round, (void *)&ng, (void *)&nx};
  return hs_host::launch_persistent(k, args, kind, num_wgs, kWg, 0, s);
}

int hs_scan_u32_noalloc(hs_arena *ar, const uint32_t *d_in, uint32_t *d_out, uint64_t n, int inclusive,
                        int kind, int num_wgs, void *stream) {
  if (!ar || (n && (!d_in || !d_out))) return HS_ERR_INVALID;
  if (kind == HS_BAR_COOP) {
    fprintf(stderr, "libhs: hs_scan_u32_noalloc: the cooperative-launch barrier kind is not offered\n");
    return HS_ERR_INVALID;
  }
  const void *k = inclusive ? scan_kernel<true>(kind) : scan_kernel<false>(kind);
  if (!k) return HS_ERR_INVALID;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 2);
  if (num_wgs > kNoAllocMaxWgs) {
    fprintf(stderr, "libhs: hs_scan_u32_noalloc: %d workgroups exceed the static scratch (%d at most)\n",
            num_wgs, kNoAllocMaxWgs);
    return HS_ERR_INVALID;
  }
  int dev = -1;
  HS_TRY_HIP(hipGetDevice(&dev));
  if (dev != ar->device || dev < 0 || dev >= kMaxDevices) return HS_ERR_INVALID;
  uint32_t *p;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch_dev[dev];
  }
  if (!p) {
    fprintf(stderr, "libhs: hs_scan_u32_noalloc: no scratch address for device %d\n", dev);
    return HS_ERR_NOT_INIT;
  }
  return hs_scan_u32(ar, d_in, d_out, n, inclusive, p, kind, num_wgs, stream);
}

}  // extern "C"
