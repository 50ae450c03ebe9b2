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

#include "hs_internal.hpp"

using namespace hs;

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

template <int KIND, bool INCLUSIVE>
__global__ void __launch_bounds__(kWg) k_scan_u32(ArenaView a, const uint32_t *__restrict__ in,
                                                 uint32_t *__restrict__ out, uint64_t n,
                                                 uint32_t *partials, uint64_t round_elems,
                                                 uint32_t ngroups, uint32_t nxcd) {
  GridBarrier gb;
  gb.init(a, KIND, ngroups, nxcd);
  const uint32_t W = gridDim.x, w = blockIdx.x;
  const uint64_t chunk = round_elems / W;  // multiple of kTile
  uint32_t carry = 0;                      // sum of all previous rounds
  for (uint64_t rb = 0; rb < n; rb += round_elems) {
    const uint64_t cb = rb + (uint64_t)w * chunk;
    uint64_t ce = cb + chunk;
    const uint64_t re = rb + round_elems < n ? rb + round_elems : n;
    if (ce > re) ce = re;
    uint32_t *part = partials + ((rb / round_elems) & 1u) * W;
    // phase 1: chunk sum
    uint32_t acc = 0;
    if (cb < ce) {
      const uint64_t full = cb + (ce - cb) / 4 * 4;
      const uint4 *p4 = reinterpret_cast<const uint4 *>(in + cb);
      const uint64_t n4 = (full - cb) / 4;
      uint64_t i = threadIdx.x;
      for (; i + 3 * kWg < n4; i += 4 * kWg) {
        uint4 v0 = p4[i], v1 = p4[i + kWg], v2 = p4[i + 2 * kWg], v3 = p4[i + 3 * kWg];
        acc += v0.x + v0.y + v0.z + v0.w + v1.x + v1.y + v1.z + v1.w;
        acc += v2.x + v2.y + v2.z + v2.w + v3.x + v3.y + v3.z + v3.w;
      }
      for (; i < n4; i += kWg) {
        uint4 v = p4[i];
        acc += v.x + v.y + v.z + v.w;
      }
      for (uint64_t j = full + threadIdx.x; j < ce; j += kWg) acc += in[j];
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[w] = acc;
    if (gb.sync<KIND>()) return;
    // phase 2: carry-in of this chunk and the round total
    uint32_t before = 0, all = 0;
    for (uint32_t j = threadIdx.x; j < W; j += kWg) {
      const uint32_t v = part[j];
      all += v;
      before += j < w ? v : 0u;
    }
    before = block_sum(before);
    all = block_sum(all);
    uint32_t run = carry + before;
    // phase 3: tile scan
    for (uint64_t tb = cb; tb < ce; tb += kTile) {
      const uint64_t base = tb + (uint64_t)threadIdx.x * kItems;
      uint32_t x[kItems];
      if (base + kItems <= ce) {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int k = 0; k < kItems / 4; ++k) {
          const uint4 v = p4[k];
          x[4 * k] = v.x; x[4 * k + 1] = v.y; x[4 * k + 2] = v.z; x[4 * k + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k) x[k] = base + k < ce ? in[base + k] : 0u;
      }
      uint32_t t = 0;
#pragma unroll
      for (int k = 0; k < kItems; ++k) t += x[k];
      uint32_t tile_total;
      uint32_t pre = run + block_excl_scan(t, tile_total);
      uint32_t y[kItems];
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        if (INCLUSIVE) { pre += x[k]; y[k] = pre; }
        else { y[k] = pre; pre += x[k]; }
      }
      if (base + kItems <= ce) {
        uint4 *o4 = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int k = 0; k < kItems / 4; ++k)
          o4[k] = make_uint4(y[4 * k], y[4 * k + 1], y[4 * k + 2], y[4 * k + 3]);
      } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
          if (base + k < ce) out[base + k] = y[k];
      }
      run += tile_total;
    }
    carry += all;
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

// round size: ~96 MiB of input per round (Infinity-Cache resident), rounded to a
// whole number of tiles per workgroup
static uint64_t scan_round_elems(int num_wgs, uint64_t n, uint64_t target_bytes) {
  const uint64_t per_wg_target = target_bytes / 4 / (uint64_t)num_wgs;
  uint64_t tiles = per_wg_target / kTile;
  if (tiles < 1) tiles = 1;
  // no more tiles per workgroup than the whole input needs
  const uint64_t need = (n + (uint64_t)num_wgs * kTile - 1) / ((uint64_t)num_wgs * kTile);
  if (tiles > need && need > 0) tiles = need;
  return tiles * kTile * (uint64_t)num_wgs;
}

static uint64_t g_scan_round_bytes = 96ull << 20;

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
  uint32_t ng = (uint32_t)v.num_groups, nx = (uint32_t)hs_host::xcd_groups();
  void *args[] = {(void *)&v, (void *)&d_in, (void *)&d_out, (void *)&n, (void *)&d_scratch,
                  (void *)&round, (void *)&ng, (void *)&nx};
  return hs_host::launch_persistent(k, args, kind, num_wgs, kWg, 0, s);
}

}  // extern "C"
