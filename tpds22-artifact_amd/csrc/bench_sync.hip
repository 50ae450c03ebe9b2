// bench_sync.hip -- the synchronization microbenchmarks: barriers, mutexes and
// semaphores over the lock arena, with per-episode invariant counters.
//
// Restates the kernel classes of Heterosync/main.cu:111-1493 (data pattern
// accessData, Heterosync/main.cu:68-108) for wave64 workgroups of any size, and
// the dispatch by syncPrim id of Heterosync/main.cu:2481-2532 / 2821-2944.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "hs_internal.hpp"

namespace cg = cooperative_groups;
using namespace hs;

namespace {

constexpr int kLineWords = 16;  // NUM_WORDS_PER_CACHELINE of the reference data pattern

struct BenchArgs {
  float *storage;
  uint32_t *chk;     // [0] violations, [64*(1+g)] per-instance in-CS counters; may be null
  int num_ldst;
  int iters;
  int ngroups;       // logical groups ("SMs") for two-level / tree / LF / local exchange
  int nxcd;          // groups of the XCD barriers
  int local_exch;
  int check;
};

// accessData (Heterosync/main.cu:68-80): thread t of data block `blk` shifts its
// (num_ldst+1)-line segment up by one line.  All loads are issued before the
// stores; the result equals the reference's descending copy because line k is
// read before line k is overwritten in both orders.
__device__ __forceinline__ void access_data(float *s, int blk, int L) {
  const int tid = blk * (int)blockDim.x + (int)threadIdx.x;
  const long base = (long)(tid / kLineWords) * (L + 1);
  const int off = threadIdx.x % kLineWords;
  constexpr int kChunk = 16;
  for (int hi = L - 1; hi >= 0; hi -= kChunk) {
    const int lo = hi - kChunk + 1 > 0 ? hi - kChunk + 1 : 0;
    float v[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      if (lo + k <= hi) v[k] = s[(base + lo + k) * kLineWords + off];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      if (lo + k <= hi) s[(base + lo + k + 1) * kLineWords + off] = v[k];
  }
}

// accessData_semRd (Heterosync/main.cu:87-98)
__device__ __forceinline__ void access_data_rd(const float *s, volatile float *dummy, int blk, int L) {
  const int tid = blk * (int)blockDim.x + (int)threadIdx.x;
  const long base = (long)(tid / kLineWords) * (L + 1);
  const int off = threadIdx.x % kLineWords;
  for (int n = L - 1; n >= 0; --n) dummy[threadIdx.x] += s[(base + n) * kLineWords + off];
}

__device__ __forceinline__ void bump(uint32_t *p) {
  if (p) add_relaxed(p, 1u);
}

// Barrier microbenchmark: kernelAtomicTreeBarrierUniq* / kernelFBSTreeBarrierUniq* /
// kernelCCGBarrierUniq* (Heterosync/main.cu:111-667).  Each episode: accessData on
// the current data block, [local barrier + exchange], device barrier, rotate.
// With p.check, each workgroup publishes the episode number before arriving and
// afterwards checks a rotating peer: a smaller value means an early pass.
template <int KIND>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_barrier_bench(ArenaView a, BenchArgs p) {
  GridBarrier gb;
  gb.init(a, KIND, (uint32_t)p.ngroups, (uint32_t)p.nxcd);
  int cur = blockIdx.x;
  for (int i = 0; i < p.iters; ++i) {
    const uint32_t e = (uint32_t)i + 1;
    access_data(p.storage, cur, p.num_ldst);
    if (p.local_exch && gb.per > 1) {
      if (gb.sync_local<KIND>()) return;
      cur = (cur + (int)gb.G) % (int)gb.nwg;
      access_data(p.storage, cur, p.num_ldst);
    }
    if (p.check && is_master()) st_relaxed(a.episode + gb.wg, e);
    if (gb.sync<KIND>()) return;
    if (p.check && is_master() && gb.nwg > 1) {
      const uint32_t peer = (gb.wg + 1 + (uint32_t)i % (gb.nwg - 1)) % gb.nwg;
      if (ld_relaxed(a.episode + peer) < e) bump(p.chk);
    }
    cur = (cur + 1) % (int)gb.nwg;
  }
}

// ---- mutex microbenchmarks (Heterosync/main.cu:669-881) ----
// Written against the public collective API (hs::mutex_lock / mutex_unlock and
// their _local forms), as the reference's kernels call cudaMutex*Lock/Unlock.
template <int MKIND, bool UNIQ>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_mutex_bench(ArenaView a, unsigned m, BenchArgs p) {
  const uint32_t g = UNIQ ? blockIdx.x % (uint32_t)p.ngroups : 0u;
  // non-uniq: all workgroups access block 0; uniq: the group's block
  const int blk = UNIQ ? (int)g : 0;
  uint32_t *inside = p.chk ? p.chk + 64 * (1 + g) : nullptr;
  for (int i = 0; i < p.iters; ++i) {
    uint32_t slot = 0;
    const int rc = UNIQ ? mutex_lock_local<MKIND>(a, m, g, &slot) : mutex_lock<MKIND>(a, m, &slot);
    if (rc) return;
    if (is_master() && inside && add_relaxed(inside, 1u) != 0u) bump(p.chk);
    access_data(p.storage, blk, p.num_ldst);
    __syncthreads();
    if (is_master() && inside) sub_release(inside, 1u);
    if (UNIQ) mutex_unlock_local<MKIND>(a, m, g, slot);
    else mutex_unlock<MKIND>(a, m, slot);
  }
}

// ---- semaphore microbenchmarks (Heterosync/main.cu:885-1493) ----
// in-CS tracking word: writers add 1<<16, readers add 1
__device__ __forceinline__ void sem_enter_check(uint32_t *chk, uint32_t *inside, bool writer,
                                                uint32_t capacity) {
  if (!inside) return;
  const uint32_t old = add_relaxed(inside, writer ? (1u << 16) : 1u);
  if (writer ? old != 0u : ((old >> 16) != 0u || (old & 0xFFFFu) + 1u > capacity)) bump(chk);
}

// non-unique: one writer per logical SM writes the blocks of every reader of that
// SM, all SMs share the same blocks (kernelSpinLockSemaphore, main.cu:885-962);
// unique: per-SM semaphore and blocks (kernelSpinLockSemaphoreUniq, 1041-1110).
// Written against hs::sem_wait / sem_post and their _local forms.
template <int SKIND, bool UNIQ>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_sem_bench(ArenaView a, unsigned sem, BenchArgs p) {
  const uint32_t nwg = gridDim.x;
  const uint32_t num_sm = (uint32_t)p.ngroups < nwg ? (uint32_t)p.ngroups : nwg;
  const uint32_t sm = UNIQ ? blockIdx.x % num_sm : blockIdx.x % (uint32_t)p.ngroups;
  const uint32_t per_sm = (nwg + num_sm - 1) / num_sm;
  const uint32_t per_sm_id = blockIdx.x / num_sm;
  const bool writer = per_sm_id == (sm % per_sm);
  const uint32_t g = UNIQ ? sm : 0u;
  uint32_t *inside = p.chk ? p.chk + 64 * (1 + g) : nullptr;
  const uint32_t capacity = ld_relaxed(sem_rec(a, sem, g) + kSemMax) - 1u;
  __shared__ volatile float dummy[1024];
  dummy[threadIdx.x] = 0.f;
  for (int i = 0; i < p.iters; ++i) {
    const int rc = UNIQ ? sem_wait_local<SKIND>(a, sem, writer, g) : sem_wait<SKIND>(a, sem, writer);
    if (rc) return;
    if (is_master()) sem_enter_check(p.chk, inside, writer, capacity);
    if (writer) {
      if (UNIQ) {
        int cur = blockIdx.x;
        for (uint32_t j = 0; j < per_sm; ++j) {
          access_data(p.storage, cur, p.num_ldst);
          cur = (cur + (int)num_sm) % (int)nwg;
        }
      } else {
        for (uint32_t j = 0; j < per_sm; ++j) access_data(p.storage, (int)j, p.num_ldst);
      }
    } else {
      access_data_rd(p.storage, dummy, UNIQ ? (int)blockIdx.x : (int)per_sm_id, p.num_ldst);
    }
    __syncthreads();
    if (is_master() && inside) sub_release(inside, writer ? (1u << 16) : 1u);
    if ((UNIQ ? sem_post_local<SKIND>(a, sem, writer, g) : sem_post<SKIND>(a, sem, writer))) return;
  }
}

// wait statistics of one acquisition (10 ns ticks of s_memrealtime): chk[1] the
// longest wait (saturating), chk[2..3] the sum of the waits (u64)
__device__ __forceinline__ void record_wait(uint32_t *chk, uint64_t ticks) {
  const uint32_t t32 = ticks > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)ticks;
  __hip_atomic_fetch_max(chk + 1, t32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(chk + 2), (unsigned long long)ticks,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Mixed reader/writer semaphore benchmark of the headline configuration: writers
// are the workgroups with wg % writer_every == 0; all share one semaphore and
// one data region (writers shift it, readers read it).  With a check buffer,
// also the in-section invariants and every acquisition's wait.
template <int SKIND>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_sem_mix(ArenaView a, unsigned sem, BenchArgs p, int writer_every) {
  const bool writer = writer_every > 0 && (blockIdx.x % (uint32_t)writer_every) == 0;
  uint32_t *inside = p.chk ? p.chk + 64 : nullptr;
  const uint32_t capacity = ld_relaxed(sem_rec(a, sem, 0) + kSemMax) - 1u;
  __shared__ volatile float dummy[1024];
  dummy[threadIdx.x] = 0.f;
  for (int i = 0; i < p.iters; ++i) {
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (sem_wait<SKIND>(a, sem, writer)) return;
    if (is_master() && inside) {
      record_wait(p.chk, __builtin_amdgcn_s_memrealtime() - t0);
      sem_enter_check(p.chk, inside, writer, capacity);
    }
    if (writer) access_data(p.storage, 0, p.num_ldst);
    else access_data_rd(p.storage, dummy, 0, p.num_ldst);
    __syncthreads();
    if (is_master() && inside) sub_release(inside, writer ? (1u << 16) : 1u);
    if (sem_post<SKIND>(a, sem, writer)) return;
  }
}

__global__ void k_zero(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int zero_async(uint32_t *p, size_t n, hipStream_t s) {
  if (!p || !n) return HS_OK;
  const int blocks = (int)((n + 255) / 256 > 1024 ? 1024 : (n + 255) / 256);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, p, n);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

const void *barrier_kernel(int kind) {
  switch (kind) {
    case HS_BAR_FLAT: return (const void *)k_barrier_bench<HS_BAR_FLAT>;
    case HS_BAR_TWO_LEVEL: return (const void *)k_barrier_bench<HS_BAR_TWO_LEVEL>;
    case HS_BAR_XCD: return (const void *)k_barrier_bench<HS_BAR_XCD>;
    case HS_BAR_XCD_FLAT_RELEASE: return (const void *)k_barrier_bench<HS_BAR_XCD_FLAT_RELEASE>;
    case HS_BAR_TREE: return (const void *)k_barrier_bench<HS_BAR_TREE>;
    case HS_BAR_LF: return (const void *)k_barrier_bench<HS_BAR_LF>;
    case HS_BAR_COOP: return (const void *)k_barrier_bench<HS_BAR_COOP>;
    case HS_BAR_NAIVE_ALL: return (const void *)k_barrier_bench<HS_BAR_NAIVE_ALL>;
    default: return nullptr;
  }
}

int launch_barrier(hs_arena *ar, int kind, const BenchArgs &p, int num_wgs, int wg_size, hipStream_t s) {
  const void *k = barrier_kernel(kind);
  if (!k) return HS_ERR_INVALID;
  int rc = hs_host::check_residency(k, num_wgs, wg_size, 0);
  if (rc) return rc;
  if (kind == HS_BAR_LF) {
    // episode numbers restart at 1 every launch
    if ((rc = zero_async(ar->view.lf_in, ar->view.max_blocks, s))) return rc;
    if ((rc = zero_async(ar->view.lf_out, ar->view.max_blocks, s))) return rc;
    if ((rc = zero_async(ar->view.lf_gin, (size_t)ar->view.num_groups * ar->view.max_blocks, s))) return rc;
    if ((rc = zero_async(ar->view.lf_gout, (size_t)ar->view.num_groups * ar->view.max_blocks, s))) return rc;
  }
  if (p.check && (rc = zero_async(ar->view.episode, ar->view.max_blocks, s))) return rc;
  hs::ArenaView v = ar->view;
  BenchArgs pa = p;
  if (kind == HS_BAR_COOP) {
    void *args[] = {(void *)&v, (void *)&pa};
    HS_TRY_HIP(hipLaunchCooperativeKernel(k, dim3(num_wgs), dim3(wg_size), args, 0, s));
  } else {
    void *args[] = {(void *)&v, (void *)&pa};
    HS_TRY_HIP(hipLaunchKernel(k, dim3(num_wgs), dim3(wg_size), args, 0, s));
  }
  return HS_OK;
}

template <int MKIND>
const void *mutex_kernel(bool uniq) {
  return uniq ? (const void *)k_mutex_bench<MKIND, true> : (const void *)k_mutex_bench<MKIND, false>;
}
template <int SKIND>
const void *sem_kernel(bool uniq) {
  return uniq ? (const void *)k_sem_bench<SKIND, true> : (const void *)k_sem_bench<SKIND, false>;
}

int launch_plain(const void *k, void **args, int num_wgs, int wg_size, hipStream_t s) {
  int rc = hs_host::check_residency(k, num_wgs, wg_size, 0);
  if (rc) return rc;
  HS_TRY_HIP(hipLaunchKernel(k, dim3(num_wgs), dim3(wg_size), args, 0, s));
  return HS_OK;
}

}  // namespace

extern "C" {

int hs_max_resident_wgs(int kind, int wg_size, int *out) {
  const void *k = barrier_kernel(kind);
  if (!k || !out || wg_size <= 0 || wg_size > 1024) return HS_ERR_INVALID;
  int per_cu = 0;
  HS_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, wg_size, 0));
  *out = per_cu * hs_host::device_cus();
  return HS_OK;
}

int hs_barrier_bench(hs_arena *ar, int kind, float *d_storage, int num_wgs, int wg_size, int num_ldst,
                     int episodes, int check, uint32_t *d_violations, void *stream) {
  if (!ar || !d_storage || num_wgs <= 0 || num_wgs > ar->view.max_blocks || wg_size <= 0 ||
      wg_size > 1024 || wg_size % 64 || num_ldst < 0 || episodes < 0)
    return HS_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  BenchArgs p{};
  p.storage = d_storage;
  p.chk = d_violations;
  p.num_ldst = num_ldst;
  p.iters = episodes;
  p.ngroups = ar->view.num_groups;
  p.nxcd = hs_host::xcd_groups(num_wgs);
  p.local_exch = 0;
  p.check = check;
  int rc;
  if (d_violations && (rc = zero_async(d_violations, 64, s))) return rc;
  return launch_barrier(ar, kind, p, num_wgs, wg_size, s);
}

// syncPrim dispatch (Heterosync/main.cu:2821-2944)
int hs_sync_bench(hs_arena *ar, int prim, float *d_storage, int num_tbs, int wg_size, int num_ldst,
                  int iterations, int check, uint32_t *d_violations, void *stream) {
  if (!ar || !d_storage || num_tbs <= 0 || num_tbs > ar->view.max_blocks || wg_size <= 0 ||
      wg_size > 1024 || num_ldst < 0 || iterations < 0)
    return HS_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int G = ar->view.num_groups;
  BenchArgs p{};
  p.storage = d_storage;
  p.chk = d_violations;
  p.num_ldst = num_ldst;
  p.iters = iterations;
  p.ngroups = G;
  p.nxcd = hs_host::xcd_groups(num_tbs);
  p.check = check;
  int rc;
  if (d_violations && (rc = zero_async(d_violations, (size_t)64 * (G + 1), s))) return rc;
  const int per_sm = (num_tbs + G - 1) / G;
  auto bar = [&](int kind, int lx) {
    p.local_exch = lx;
    return launch_barrier(ar, kind, p, num_tbs, wg_size, s);
  };
  hs::ArenaView v = ar->view;
  switch (prim) {
    case 0: return bar(HS_BAR_TREE, 0);
    case 1: return bar(HS_BAR_TREE, 1);
    case 2: return bar(HS_BAR_LF, 0);
    case 3: return bar(HS_BAR_LF, 1);
    case 37: return bar(HS_BAR_TWO_LEVEL, 0);
    case 38: return bar(HS_BAR_TWO_LEVEL, 1);
    case 39: return bar(HS_BAR_FLAT, 0);
    case 40: return bar(HS_BAR_FLAT, 1);
    case 41: return bar(HS_BAR_NAIVE_ALL, 0);
    case 42: return bar(HS_BAR_NAIVE_ALL, 1);
    // hybrid: CCG when TBs/SM < (MAX_BLOCKS/NUM_SM)/4 (Heterosync/main.cu:1595-1610)
    case 43: return bar(per_sm < (ar->view.max_blocks / G) / 4 ? HS_BAR_COOP : HS_BAR_TWO_LEVEL, 0);
    case 44: return bar(per_sm < (ar->view.max_blocks / G) / 4 ? HS_BAR_COOP : HS_BAR_TWO_LEVEL, 1);
    case 45: return bar(HS_BAR_COOP, 0);
    case 46: return bar(HS_BAR_COOP, 1);
    default: break;
  }
  // mutexes: ids per Heterosync/main.cu:2638-2700
  if ((prim >= 4 && prim <= 7) || (prim >= 20 && prim <= 23)) {
    const bool uniq = prim >= 20;
    const int kind = (prim - (uniq ? 20 : 4));  // spin, EBO, sleep, FA order of the ids
    const unsigned m = (unsigned)(prim - 4 >= 16 ? prim - 16 : prim - 4);  // mutex numbers 0..7
    if ((int)m >= v.num_mutexes) return HS_ERR_INVALID;
    void *args[] = {(void *)&v, (void *)&m, (void *)&p};
    const void *k = nullptr;
    switch (kind) {
      case 0: k = mutex_kernel<HS_MUTEX_SPIN>(uniq); break;
      case 1: k = mutex_kernel<HS_MUTEX_EBO>(uniq); break;
      case 2: k = mutex_kernel<HS_MUTEX_SLEEP>(uniq); break;
      case 3: k = mutex_kernel<HS_MUTEX_FA>(uniq); break;
    }
    return launch_plain(k, args, num_tbs, wg_size, s);
  }
  // semaphores: ids per Heterosync/main.cu:2701-2780; semaphore numbers and
  // capacities follow the reference's create calls
  if ((prim >= 8 && prim <= 19) || (prim >= 24 && prim <= 35)) {
    static const unsigned caps[4] = {1, 2, 10, 120};
    const int idx = (prim - 8) % 4;
    unsigned sem;
    int skind;
    bool uniq = false;
    if (prim <= 11) { sem = (unsigned)idx; skind = HS_SEM_SPIN; }
    else if (prim <= 15) { sem = 4u + idx; skind = HS_SEM_EBO; }
    else if (prim <= 19) { sem = (unsigned)idx; skind = HS_SEM_SPIN_PRIORITY; }
    else if (prim <= 27) { sem = 12u + idx; skind = HS_SEM_SPIN; uniq = true; }
    else if (prim <= 31) { sem = 16u + idx; skind = HS_SEM_EBO; uniq = true; }
    else { sem = 4u + idx; skind = HS_SEM_EBO_PRIORITY; }
    if ((int)sem >= v.num_sems) return HS_ERR_INVALID;
    unsigned h = 0;
    if ((rc = hs_sem_create(ar, skind, sem, caps[idx], &h))) return rc;
    void *args[] = {(void *)&v, (void *)&sem, (void *)&p};
    const void *k = nullptr;
    switch (skind) {
      case HS_SEM_SPIN: k = sem_kernel<HS_SEM_SPIN>(uniq); break;
      case HS_SEM_EBO: k = sem_kernel<HS_SEM_EBO>(uniq); break;
      case HS_SEM_SPIN_PRIORITY: k = sem_kernel<HS_SEM_SPIN_PRIORITY>(uniq); break;
      case HS_SEM_EBO_PRIORITY: k = sem_kernel<HS_SEM_EBO_PRIORITY>(uniq); break;
    }
    return launch_plain(k, args, num_tbs, wg_size, s);
  }
  return HS_ERR_INVALID;
}

int hs_sem_bench(hs_arena *ar, int kind, unsigned sem, float *d_storage, int num_wgs, int wg_size,
                 int num_ldst, int iterations, int writer_every, uint32_t *d_violations, void *stream) {
  if (!ar || !d_storage || num_wgs <= 0 || num_wgs > ar->view.max_blocks || wg_size <= 0 ||
      wg_size > 1024 || (int)sem >= ar->view.num_sems || kind < HS_SEM_SPIN || kind > HS_SEM_XCD_QUEUE)
    return HS_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  BenchArgs p{};
  p.storage = d_storage;
  p.chk = d_violations;
  p.num_ldst = num_ldst;
  p.iters = iterations;
  p.ngroups = ar->view.num_groups;
  int rc;
  if (d_violations && (rc = zero_async(d_violations, 128, s))) return rc;
  hs::ArenaView v = ar->view;
  void *args[] = {(void *)&v, (void *)&sem, (void *)&p, (void *)&writer_every};
  const void *k = nullptr;
  switch (kind) {
    case HS_SEM_SPIN: k = (const void *)k_sem_mix<HS_SEM_SPIN>; break;
    case HS_SEM_EBO: k = (const void *)k_sem_mix<HS_SEM_EBO>; break;
    case HS_SEM_SPIN_PRIORITY: k = (const void *)k_sem_mix<HS_SEM_SPIN_PRIORITY>; break;
    case HS_SEM_EBO_PRIORITY: k = (const void *)k_sem_mix<HS_SEM_EBO_PRIORITY>; break;
    case HS_SEM_XCD_QUEUE: k = (const void *)k_sem_mix<HS_SEM_XCD_QUEUE>; break;
  }
  return launch_plain(k, args, num_wgs, wg_size, s);
}

}  // extern "C"
