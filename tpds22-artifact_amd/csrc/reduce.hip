// reduce.hip -- persistent int32 sum reduction with one grid barrier.
//
// Replaces reduce_kernel / reduce() (Reduce/reduce_gsrb.cu:289-404,
// Reduce/reduce_ccg.cu:9-88).  The reference launches ceil(N/tpb) blocks and
// runs log_tpb(N) levels separated by barriers, which cannot be co-resident at
// N = 2^24.  Here a co-resident grid (default 1 workgroup per CU x 256 threads:
// measured faster than 2 per CU, fewer arrivals for the same streaming rate)
// streams the input once with 16-byte loads (8 in flight per thread), reduces
// each wave with DPP row operations, combines the 4 waves in LDS, and crosses
// ONE grid episode: with the XCD barriers its arrival half only, the
// workgroup sums riding on the arrival atomics (GridBarrier::arrive_sum; the
// last arriver holds the total), with the other kinds a full episode after
// which workgroup 0 folds the published partials.  Arithmetic is
// uint32 (two's-complement wrap, as the reference's int sum), so every grid
// size and schedule gives the same bits.
#include <hip/hip_runtime.h>

#include "hs_internal.hpp"

using namespace hs;

namespace {

constexpr int kWg = 256;
constexpr int kUnroll = 8;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// wave64 sum: DPP row_shr within rows of 16, then row_bcast15 / row_bcast31
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
  __shared__ uint32_t part[kWg / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int i = 0; i < kWg / 64; ++i) t += part[i];
  __syncthreads();
  return t;
}

// STAMPS (the diagnostic form, hs_reduce_i32_stamped): thread 0 of workgroup b
// writes stamps[4b..4b+2] = the device's 100 MHz clock when the workgroup
// starts, when its loads are summed, and when it leaves (after the fold for the
// last arriver), so the host can split a launch into dispatch spread,
// streaming and fold.
template <int KIND, bool STAMPS = false>
__global__ void __launch_bounds__(kWg) k_reduce_i32(ArenaView a, const int32_t *__restrict__ in,
                                                   uint64_t n, int32_t *partials, int32_t *out,
                                                   uint32_t ngroups, uint32_t nxcd, uint64_t *stamps) {
  uint64_t t_start = 0;
  if (STAMPS && threadIdx.x == 0) t_start = __builtin_amdgcn_s_memrealtime();
  GridBarrier gb;
  gb.init(a, KIND, ngroups, nxcd);
  const uint64_t n4 = n / 4;
  const u32x4 *in4 = reinterpret_cast<const u32x4 *>(in);
  const uint64_t stride = (uint64_t)gridDim.x * kWg;
  uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
    u32x4 v[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) v[k] = __builtin_nontemporal_load(in4 + i + k * stride);
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) acc += v[k].x + v[k].y + v[k].z + v[k].w;
  }
  for (; i < n4; i += stride) {
    const u32x4 v = __builtin_nontemporal_load(in4 + i);
    acc += v.x + v.y + v.z + v.w;
  }
  // tail (n % 4 elements)
  const uint64_t t = n4 * 4 + (uint64_t)blockIdx.x * kWg + threadIdx.x;
  if (t < n && (uint64_t)blockIdx.x * kWg + threadIdx.x < 4) acc += (uint32_t)in[t];
  acc = block_sum(acc);
  if (STAMPS && threadIdx.x == 0) {
    stamps[4 * blockIdx.x + 0] = t_start;
    stamps[4 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
  }
  if constexpr (KIND == HS_BAR_XCD || KIND == HS_BAR_XCD_FLAT_RELEASE) {
    // the XCD kinds fold inside the fan-in: each workgroup's sum rides on its
    // arrival atomics, and the last arriver holds the total
    uint32_t total;
    if (gb.arrive_sum<KIND>(acc, &total) && threadIdx.x == 0) out[0] = (int32_t)total;
    if (STAMPS && threadIdx.x == 0) stamps[4 * blockIdx.x + 2] = __builtin_amdgcn_s_memrealtime();
    return;
  }
  // partials travel with agent-scope stores and loads; the other kinds run a
  // full episode and workgroup 0 folds
  if (threadIdx.x == 0) st_relaxed(reinterpret_cast<uint32_t *>(partials) + blockIdx.x, acc);
  if (gb.arrive_last<KIND>() <= 0) {
    if (STAMPS && threadIdx.x == 0) stamps[4 * blockIdx.x + 2] = __builtin_amdgcn_s_memrealtime();
    return;
  }
  uint32_t s = 0;
  for (uint32_t w = threadIdx.x; w < gridDim.x; w += kWg)
    s += ld_relaxed(reinterpret_cast<const uint32_t *>(partials) + w);
  s = block_sum(s);
  if (threadIdx.x == 0) out[0] = (int32_t)s;
  if (STAMPS && threadIdx.x == 0) stamps[4 * blockIdx.x + 2] = __builtin_amdgcn_s_memrealtime();
}

const void *reduce_kernel(int kind) {
  switch (kind) {
    case HS_BAR_FLAT: return (const void *)k_reduce_i32<HS_BAR_FLAT>;
    case HS_BAR_TWO_LEVEL: return (const void *)k_reduce_i32<HS_BAR_TWO_LEVEL>;
    case HS_BAR_XCD: return (const void *)k_reduce_i32<HS_BAR_XCD>;
    case HS_BAR_XCD_FLAT_RELEASE: return (const void *)k_reduce_i32<HS_BAR_XCD_FLAT_RELEASE>;
    case HS_BAR_TREE: return (const void *)k_reduce_i32<HS_BAR_TREE>;
    case HS_BAR_LF: return (const void *)k_reduce_i32<HS_BAR_LF>;
    case HS_BAR_COOP: return (const void *)k_reduce_i32<HS_BAR_COOP>;
    case HS_BAR_NAIVE_ALL: return (const void *)k_reduce_i32<HS_BAR_NAIVE_ALL>;
    default: return nullptr;
  }
}

const void *reduce_kernel_stamped(int kind) {
  switch (kind) {
    case HS_BAR_FLAT: return (const void *)k_reduce_i32<HS_BAR_FLAT, true>;
    case HS_BAR_TWO_LEVEL: return (const void *)k_reduce_i32<HS_BAR_TWO_LEVEL, true>;
    case HS_BAR_XCD: return (const void *)k_reduce_i32<HS_BAR_XCD, true>;
    case HS_BAR_COOP: return (const void *)k_reduce_i32<HS_BAR_COOP, true>;
    default: return nullptr;
  }
}

__global__ void k_zero_words(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace

namespace hs_host {
// zero the lock-free barrier flags a persistent launch will use (fill kernels,
// capture-safe).  The device status word is not cleared per launch: a failure
// stays visible (and makes later waits give up) until hs_arena_sync reports it
// and resets the arena.
int prepare_barrier(hs_arena *ar, int kind, hipStream_t s) {
  if (kind == HS_BAR_LF) {
    const hs::ArenaView &v = ar->view;
    hipLaunchKernelGGL(k_zero_words, dim3(8), dim3(256), 0, s, v.lf_in, (size_t)v.max_blocks);
    hipLaunchKernelGGL(k_zero_words, dim3(8), dim3(256), 0, s, v.lf_out, (size_t)v.max_blocks);
    hipLaunchKernelGGL(k_zero_words, dim3(256), dim3(256), 0, s, v.lf_gin,
                       (size_t)v.num_groups * v.max_blocks);
    hipLaunchKernelGGL(k_zero_words, dim3(256), dim3(256), 0, s, v.lf_gout,
                       (size_t)v.num_groups * v.max_blocks);
    HS_TRY_HIP(hipGetLastError());
  }
  return HS_OK;
}

int launch_persistent(const void *k, void **args, int kind, int num_wgs, int wg_size, size_t lds,
                      hipStream_t s) {
  if (kind == HS_BAR_COOP) {
    HS_TRY_HIP(hipLaunchCooperativeKernel(k, dim3(num_wgs), dim3(wg_size), args, (unsigned)lds, s));
  } else {
    HS_TRY_HIP(hipLaunchKernel(k, dim3(num_wgs), dim3(wg_size), args, lds, s));
  }
  return HS_OK;
}
}  // namespace hs_host

extern "C" {

int hs_reduce_partials_size(int num_wgs, uint64_t n, uint64_t *elems_out) {
  (void)n;
  if (!elems_out) return HS_ERR_INVALID;
  if (num_wgs <= 0) num_wgs = hs_host::device_cus();
  *elems_out = (uint64_t)num_wgs;
  return HS_OK;
}

static int reduce_launch(hs_arena *ar, const int32_t *d_in, uint64_t n, int32_t *d_partials, int32_t *d_out,
                         int kind, int num_wgs, uint64_t *d_stamps, void *stream) {
  if (!ar || !d_out || !d_partials || (n && !d_in)) return HS_ERR_INVALID;
  if (((uintptr_t)d_in & 15u) != 0) return HS_ERR_INVALID;  // 16-byte loads
  const void *k = d_stamps ? reduce_kernel_stamped(kind) : reduce_kernel(kind);
  if (!k) return HS_ERR_INVALID;
  if (num_wgs <= 0) num_wgs = hs_host::default_grid(k, kWg, 1);
  if (num_wgs > ar->view.max_blocks) return HS_ERR_INVALID;
  int rc = hs_host::check_residency(k, num_wgs, kWg, 0);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = hs_host::prepare_barrier(ar, kind, s))) return rc;
  hs::ArenaView v = ar->view;
  uint32_t ng = (uint32_t)v.num_groups, nx = (uint32_t)hs_host::xcd_groups(num_wgs);
  void *args[] = {(void *)&v, (void *)&d_in, (void *)&n, (void *)&d_partials, (void *)&d_out,
                  (void *)&ng, (void *)&nx, (void *)&d_stamps};
  return hs_host::launch_persistent(k, args, kind, num_wgs, kWg, 0, s);
}

int hs_reduce_i32(hs_arena *ar, const int32_t *d_in, uint64_t n, int32_t *d_partials, int32_t *d_out,
                  int kind, int num_wgs, void *stream) {
  return reduce_launch(ar, d_in, n, d_partials, d_out, kind, num_wgs, nullptr, stream);
}

int hs_reduce_i32_stamped(hs_arena *ar, const int32_t *d_in, uint64_t n, int32_t *d_partials, int32_t *d_out,
                          int kind, int num_wgs, uint64_t *d_stamps, void *stream) {
  if (!d_stamps) return HS_ERR_INVALID;
  return reduce_launch(ar, d_in, n, d_partials, d_out, kind, num_wgs, d_stamps, stream);
}

}  // extern "C"
