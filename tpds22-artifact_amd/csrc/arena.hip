// arena.hip -- lock arena, handles, singleton API and launch helpers.
//
// Replaces cudaLocksInit / cudaLocksDestroy (Heterosync/cudaLocks.cu:3-102) and the
// per-app managed-memory barrier blocks (Reduce/reduce_gsrb.cu:345-374).  One
// device allocation holds every region; hot words sit 256 B apart so that
// independent counters never share a cache line; initialisation is a handful of
// fill kernels instead of one cudaMemset per int.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string.h>

#include "hs_internal.hpp"

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ring buffers: slot 0 of every (mutex, group) ring = 1, others = -1
// (Heterosync/cudaLocks.cu:50-63)
__global__ void k_init_rings(int32_t *rings, size_t n, int stride) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    rings[i] = (i % (size_t)stride) == 0 ? 1 : -1;
}

// semaphore record of every group: count = max = capacity + 1 (the "+1" lets the
// count double as lock, Heterosync/cudaLocksSemaphoreSpin.cu:11-26)
__global__ void k_init_sem(uint32_t *sems, int sem, int groups, uint32_t init) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += gridDim.x * blockDim.x) {
    uint32_t *rec = sems + (size_t)(sem * groups + g) * hs::kSemWords;
    for (int w = 0; w < hs::kSemWords; ++w) rec[w] = 0;
    rec[hs::kSemCount] = init;
    rec[hs::kSemMax] = init;
  }
}

__global__ void k_probe_xcc(int *out) {
  if (threadIdx.x == 0) out[blockIdx.x] = (int)hs::xcc_id();
}

// one wave: thread 0 polls *word until it equals value (bounded, abortable)
__global__ void k_wait_value(hs::ArenaView a, const uint32_t *word, uint32_t value) {
  if (threadIdx.x == 0)
    hs::spin_until(a, [&] { return hs::ld_relaxed(word) == value; }, 8);
}

int launch_fill(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
  if (!n) return HS_OK;
  const int blocks = (int)((n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

}  // namespace

namespace hs_host {

int device_cus() {
  static int cached = 0;
  if (!cached) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cached = p.multiProcessorCount;
    if (cached <= 0) cached = 1;
  }
  return cached;
}

int device_xcds() {
  static int cached = 0;
  if (!cached) {
    const int n = 4 * 64;
    int *d = nullptr;
    int h[4 * 64];
    cached = 1;
    if (hipMalloc(&d, sizeof(int) * n) == hipSuccess) {
      hipLaunchKernelGGL(k_probe_xcc, dim3(n), dim3(64), 0, 0, d);
      if (hipMemcpy(h, d, sizeof(int) * n, hipMemcpyDeviceToHost) == hipSuccess) {
        int mx = 0;
        for (int i = 0; i < n; ++i) mx = h[i] > mx ? h[i] : mx;
        cached = mx + 1;
      }
      (void)hipFree(d);  // a probe buffer: nothing to do if the free fails
    }
  }
  return cached;
}

static int g_xcd_groups = 0;

int xcd_groups_override() { return g_xcd_groups; }

int xcd_groups(int num_wgs) {
  if (g_xcd_groups > 0) return g_xcd_groups;
  // one group per XCD, doubled while a group would exceed 255 members: the
  // XCD barrier's per-XCD arrival bytes hold at most 255, and a multiple of the
  // XCD count keeps every group on one XCD under round-robin dispatch
  int x = device_xcds();
  while ((num_wgs + x - 1) / x > 255 && 2 * x <= hs::kMaxGroups) x *= 2;
  return x;
}

int check_residency(const void *kernel, int num_wgs, int wg_size, size_t dyn_lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wg_size, dyn_lds) != hipSuccess)
    return HS_ERR_HIP;
  if ((long long)per_cu * device_cus() < (long long)num_wgs) return HS_ERR_RESIDENCY;
  return HS_OK;
}

int default_grid(const void *kernel, int wg_size, int wgs_per_cu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wg_size, 0) != hipSuccess)
    per_cu = 1;
  if (wgs_per_cu > per_cu) wgs_per_cu = per_cu;
  if (wgs_per_cu < 1) wgs_per_cu = 1;
  return wgs_per_cu * device_cus();
}

int clear_status(hs_arena *a, hipStream_t s) {
  // a fill kernel, not hipMemsetAsync: the launchers may be captured in a graph
  return launch_fill(a->view.status, 1, 0u, s);
}

int finish(hs_arena *a, hipStream_t s) {
  HS_TRY_HIP(hipGetLastError());
  HS_TRY_HIP(hipStreamSynchronize(s));
  uint32_t st = 0;
  HS_TRY_HIP(hipMemcpy(&st, a->view.status, sizeof(st), hipMemcpyDeviceToHost));
  if (st != 0) {
    // a failed wait leaves counters half-way through an episode and semaphore
    // records half-acquired: start clean (hs_arena_reset restores every
    // semaphore to its created capacity)
    if (hs_arena_reset(a, s) != HS_OK) return HS_ERR_HIP;
    return st == hs::kErrAborted ? HS_ERR_ABORTED : HS_ERR_TIMEOUT;
  }
  return HS_OK;
}

int arena_init_words(hs_arena *a, hipStream_t s) {
  const hs::ArenaView &v = a->view;
  int rc;
  if ((rc = launch_fill(v.bar, hs::BarOff::kWords, 0u, s))) return rc;
  if ((rc = launch_fill(v.lf_in, (size_t)v.max_blocks, 0u, s))) return rc;
  if ((rc = launch_fill(v.lf_out, (size_t)v.max_blocks, 0u, s))) return rc;
  if ((rc = launch_fill(v.lf_gin, (size_t)v.num_groups * v.max_blocks, 0u, s))) return rc;
  if ((rc = launch_fill(v.lf_gout, (size_t)v.num_groups * v.max_blocks, 0u, s))) return rc;
  const size_t mwords = (size_t)v.num_mutexes * v.num_groups * hs::kPadWords;
  if ((rc = launch_fill(v.mutex_heads, mwords, 0u, s))) return rc;
  if ((rc = launch_fill(v.mutex_tails, mwords, 0u, s))) return rc;
  const size_t rn = (size_t)v.num_mutexes * v.num_groups * v.ring_stride;
  if (rn) {
    const int blocks = (int)((rn + 255) / 256 > 4096 ? 4096 : (rn + 255) / 256);
    hipLaunchKernelGGL(k_init_rings, dim3(blocks), dim3(256), 0, s, v.mutex_rings, rn, v.ring_stride);
    HS_TRY_HIP(hipGetLastError());
  }
  // semaphore records are not touched here: hs_arena_reset re-creates each one
  // with its created capacity
  if ((rc = launch_fill(v.episode, (size_t)v.max_blocks, 0u, s))) return rc;
  if ((rc = launch_fill(v.status, 1, 0u, s))) return rc;
  return HS_OK;
}

}  // namespace hs_host

// per-arena record of created semaphore capacities so that reset can restore them
namespace {
struct SemCaps {
  hs_arena *arena;
  uint32_t *caps;  // host array [num_sems], 0 = not created
};
std::mutex g_caps_mu;
SemCaps g_caps[64];

uint32_t *caps_for(hs_arena *a, bool create) {
  std::lock_guard<std::mutex> lk(g_caps_mu);
  for (auto &c : g_caps)
    if (c.arena == a) return c.caps;
  if (!create) return nullptr;
  for (auto &c : g_caps)
    if (!c.arena) {
      c.arena = a;
      c.caps = (uint32_t *)calloc((size_t)a->view.num_sems, sizeof(uint32_t));
      return c.caps;
    }
  return nullptr;
}
void caps_drop(hs_arena *a) {
  std::lock_guard<std::mutex> lk(g_caps_mu);
  for (auto &c : g_caps)
    if (c.arena == a) {
      free(c.caps);
      c.arena = nullptr;
      c.caps = nullptr;
    }
}
}  // namespace

extern "C" {

int hs_arena_create(const hs_arena_desc *desc, hs_arena **out) {
  if (!desc || !out) return HS_ERR_INVALID;
  *out = nullptr;
  // Heterosync/cudaLocks.cu:13-15: non-positive sizes are an initialisation error
  if (desc->max_blocks <= 0 || desc->num_mutexes <= 0 || desc->num_semaphores <= 0 ||
      desc->num_groups <= 0 || desc->num_groups > hs::kMaxGroups)
    return HS_ERR_INVALID;
  hs_arena *a = (hs_arena *)calloc(1, sizeof(hs_arena));
  if (!a) return HS_ERR_NOMEM;
  a->desc = *desc;
  if (hipGetDevice(&a->device) != hipSuccess) {
    free(a);
    return HS_ERR_HIP;
  }
  a->num_cus = hs_host::device_cus();
  // the XCD probe allocates and frees device memory, which waits for the whole
  // device: do it now, not first inside a launcher while a persistent loop runs
  (void)hs_host::device_xcds();

  hs::ArenaView &v = a->view;
  v.max_blocks = desc->max_blocks;
  v.num_groups = desc->num_groups;
  v.num_mutexes = desc->num_mutexes;
  v.num_sems = desc->num_semaphores;
  v.queue_len = 0;
  // ring index runs 0..max_blocks inclusive (atomicInc wrap), as arrayStride
  // = (maxBlocks + NUM_SM)/16*16 in Heterosync/cudaLocks.cu:19
  v.ring_stride = (int)align_up((size_t)desc->max_blocks + 1, 16);
  v.spin_limit = desc->spin_limit ? desc->spin_limit : hs::kDefaultSpinLimit;

  const size_t al = desc->page_align ? 4096 : 256;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, al);
    return o;
  };
  const size_t o_bar = take(sizeof(uint32_t) * hs::BarOff::kWords);
  const size_t o_lfin = take(sizeof(uint32_t) * v.max_blocks);
  const size_t o_lfout = take(sizeof(uint32_t) * v.max_blocks);
  const size_t o_lfgin = take(sizeof(uint32_t) * (size_t)v.num_groups * v.max_blocks);
  const size_t o_lfgout = take(sizeof(uint32_t) * (size_t)v.num_groups * v.max_blocks);
  const size_t mwords = (size_t)v.num_mutexes * v.num_groups * hs::kPadWords;
  const size_t o_mh = take(sizeof(uint32_t) * mwords);
  const size_t o_mt = take(sizeof(uint32_t) * mwords);
  const size_t o_ring = take(sizeof(int32_t) * (size_t)v.num_mutexes * v.num_groups * v.ring_stride);
  const size_t o_sem = take(sizeof(uint32_t) * (size_t)v.num_sems * v.num_groups * hs::kSemWords);
  const size_t o_ep = take(sizeof(uint32_t) * v.max_blocks);
  const size_t o_st = take(sizeof(uint32_t) * hs::kPadWords);
  a->bytes = off;
  if (hipMalloc(&a->base, a->bytes) != hipSuccess) {
    free(a);
    return HS_ERR_NOMEM;
  }
  char *b = (char *)a->base;
  v.bar = (uint32_t *)(b + o_bar);
  v.lf_in = (uint32_t *)(b + o_lfin);
  v.lf_out = (uint32_t *)(b + o_lfout);
  v.lf_gin = (uint32_t *)(b + o_lfgin);
  v.lf_gout = (uint32_t *)(b + o_lfgout);
  v.mutex_heads = (uint32_t *)(b + o_mh);
  v.mutex_tails = (uint32_t *)(b + o_mt);
  v.mutex_rings = (int32_t *)(b + o_ring);
  v.sems = (uint32_t *)(b + o_sem);
  v.sem_queues = nullptr;
  v.episode = (uint32_t *)(b + o_ep);
  v.status = (uint32_t *)(b + o_st);
  if (hipHostMalloc((void **)&a->abort_host, sizeof(uint32_t) * 64, hipHostMallocCoherent) !=
      hipSuccess) {
    (void)hipFree(a->base);
    free(a);
    return HS_ERR_NOMEM;
  }
  memset(a->abort_host, 0, sizeof(uint32_t) * 64);
  v.abort_flag = a->abort_host;
  int rc = hs_host::arena_init_words(a, 0);
  if (rc == HS_OK)
    rc = launch_fill(v.sems, (size_t)v.num_sems * v.num_groups * hs::kSemWords, 0u, 0);
  if (rc == HS_OK && hipDeviceSynchronize() != hipSuccess) rc = HS_ERR_HIP;
  if (rc != HS_OK) {  // already failing: the frees' own status adds nothing
    (void)hipHostFree(a->abort_host);
    (void)hipFree(a->base);
    free(a);
    return rc;
  }
  caps_for(a, true);
  // the no-allocation scan's static scratch: its address on this device, once
  // (a failure only disables that entry point, which then reports it)
  (void)hs_host::scan_scratch_init(a->device);
  *out = a;
  return HS_OK;
}

int hs_arena_destroy(hs_arena *a) {
  if (!a) return HS_ERR_INVALID;
  // work still using the arena must end before its memory goes; a failed
  // wait is reported, the memory is released either way
  const hipError_t e = hipDeviceSynchronize();
  caps_drop(a);
  const hipError_t f = hipFree(a->base);
  const hipError_t h = hipHostFree(a->abort_host);
  free(a);
  return (e == hipSuccess && f == hipSuccess && h == hipSuccess) ? HS_OK : HS_ERR_HIP;
}

int hs_arena_reset(hs_arena *a, void *stream) {
  if (!a) return HS_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::arena_init_words(a, s);
  if (rc) return rc;
  uint32_t *caps = caps_for(a, false);
  for (int i = 0; caps && i < a->view.num_sems; ++i)
    if (caps[i]) {
      hipLaunchKernelGGL(k_init_sem, dim3(1), dim3(256), 0, s, a->view.sems, i, a->view.num_groups,
                         caps[i] + 1);
    }
  HS_TRY_HIP(hipGetLastError());
  HS_TRY_HIP(hipStreamSynchronize(s));
  return HS_OK;
}

int hs_arena_set_abort(hs_arena *a, int value) {
  if (!a) return HS_ERR_INVALID;
  __atomic_store_n(a->abort_host, (uint32_t)value, __ATOMIC_SEQ_CST);
  return HS_OK;
}

int hs_wait_value_u32(hs_arena *a, const uint32_t *d_word, uint32_t value, void *stream) {
  if (!a || !d_word) return HS_ERR_INVALID;
  hipLaunchKernelGGL(k_wait_value, dim3(1), dim3(64), 0, (hipStream_t)stream, a->view, d_word, value);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

int hs_arena_device_status(hs_arena *a, int *status_out) {
  if (!a || !status_out) return HS_ERR_INVALID;
  uint32_t st = 0;
  HS_TRY_HIP(hipMemcpy(&st, a->view.status, sizeof(st), hipMemcpyDeviceToHost));
  *status_out = (int)st;
  return HS_OK;
}

int hs_arena_episode_slots(hs_arena *a, uint32_t *h_out, int count) {
  if (!a || !h_out || count <= 0 || count > a->view.max_blocks) return HS_ERR_INVALID;
  HS_TRY_HIP(hipMemcpy(h_out, a->view.episode, sizeof(uint32_t) * (size_t)count, hipMemcpyDeviceToHost));
  return HS_OK;
}

int hs_arena_sync(hs_arena *a, void *stream) {
  if (!a) return HS_ERR_INVALID;
  return hs_host::finish(a, (hipStream_t)stream);
}

int hs_arena_device_view(hs_arena *a, void *view_out, size_t view_bytes) {
  if (!a || !view_out || view_bytes != sizeof(hs::ArenaView)) return HS_ERR_INVALID;
  memcpy(view_out, &a->view, sizeof(hs::ArenaView));
  return HS_OK;
}

int hs_mutex_create(hs_arena *a, int kind, unsigned mutex_number, unsigned *handle) {
  if (!a || !handle || kind < HS_MUTEX_SPIN || kind > HS_MUTEX_SLEEP) return HS_ERR_INVALID;
  if ((int)mutex_number >= a->view.num_mutexes) return HS_ERR_INVALID;
  // mutex words start unlocked (heads/tails 0, ring slot 0 = 1): the handle is
  // the index, as in cudaMutexCreateSpin (Heterosync/cudaLocksMutexSpin.cu:4-9)
  *handle = mutex_number;
  return HS_OK;
}

int hs_sem_create(hs_arena *a, int kind, unsigned sem_number, unsigned count, unsigned *handle) {
  if (!a || !handle || kind < HS_SEM_SPIN || kind > HS_SEM_XCD_QUEUE) return HS_ERR_INVALID;
  if ((int)sem_number >= a->view.num_sems || count == 0 || count == 0xFFFFFFFFu)
    return HS_ERR_INVALID;
  hipLaunchKernelGGL(k_init_sem, dim3(1), dim3(256), 0, 0, a->view.sems, (int)sem_number,
                     a->view.num_groups, count + 1);
  HS_TRY_HIP(hipGetLastError());
  HS_TRY_HIP(hipDeviceSynchronize());
  uint32_t *caps = caps_for(a, true);
  if (caps) caps[sem_number] = count;
  *handle = sem_number;
  return HS_OK;
}

// ---- singleton API (Heterosync/cudaLocks.h:21-26) ----
static hs_arena *g_singleton = nullptr;
static int g_singleton_groups = 0;

int hsLocksInit(int maxBlocksPerKernel, int numMutexes, int numSemaphores, int pageAlign,
                int NUM_SM) {
  if (g_singleton) return HS_ERR_INVALID;
  if (maxBlocksPerKernel <= 0 || numMutexes <= 0 || numSemaphores <= 0 || NUM_SM <= 0)
    return HS_ERR_INVALID;
  // the reference sizes per-SM instances into the counts (8*NUM_SM mutexes,
  // 24*NUM_SM semaphores, Heterosync/main.cu:2542); here the group dimension is
  // explicit, so the distinct counts are the quotients
  hs_arena_desc d;
  memset(&d, 0, sizeof(d));
  d.max_blocks = maxBlocksPerKernel;
  d.num_mutexes = (numMutexes + NUM_SM - 1) / NUM_SM;
  d.num_semaphores = (numSemaphores + NUM_SM - 1) / NUM_SM;
  d.num_groups = NUM_SM;
  d.page_align = pageAlign;
  g_singleton_groups = NUM_SM;
  return hs_arena_create(&d, &g_singleton);
}

int hsLocksDestroy(void) {
  if (!g_singleton) return HS_ERR_NOT_INIT;
  int rc = hs_arena_destroy(g_singleton);
  g_singleton = nullptr;
  return rc;
}

hs_arena *hsLocksArena(void) { return g_singleton; }

#define HS_SINGLETON_MUTEX(NAME, KIND)                                  \
  int NAME(unsigned *handle, int mutexNumber) {                          \
    if (!g_singleton) return HS_ERR_NOT_INIT;                            \
    return hs_mutex_create(g_singleton, KIND, (unsigned)mutexNumber, handle); \
  }
HS_SINGLETON_MUTEX(hsMutexCreateSpin, HS_MUTEX_SPIN)
HS_SINGLETON_MUTEX(hsMutexCreateEBO, HS_MUTEX_EBO)
HS_SINGLETON_MUTEX(hsMutexCreateFA, HS_MUTEX_FA)
HS_SINGLETON_MUTEX(hsMutexCreateSleep, HS_MUTEX_SLEEP)

int hsSemaphoreCreateSpin(unsigned *handle, int semaphoreNumber, unsigned count, int NUM_SM) {
  if (!g_singleton) return HS_ERR_NOT_INIT;
  if (NUM_SM != g_singleton_groups) return HS_ERR_INVALID;
  return hs_sem_create(g_singleton, HS_SEM_SPIN, (unsigned)semaphoreNumber, count, handle);
}

int hsSemaphoreCreateEBO(unsigned *handle, int semaphoreNumber, unsigned count, int NUM_SM) {
  if (!g_singleton) return HS_ERR_NOT_INIT;
  if (NUM_SM != g_singleton_groups) return HS_ERR_INVALID;
  return hs_sem_create(g_singleton, HS_SEM_EBO, (unsigned)semaphoreNumber, count, handle);
}

int hs_set_xcd_groups(int groups) {
  if (groups < 0 || groups > hs::kMaxGroups) return HS_ERR_INVALID;
  hs_host::g_xcd_groups = groups;
  return HS_OK;
}

// ---- persistent kernels of the caller's own ----
int hs_barrier_fanin(hs_arena *ar, int num_wgs, uint32_t *ngroups, uint32_t *nxcd) {
  if (!ar || num_wgs <= 0 || num_wgs > ar->view.max_blocks || !ngroups || !nxcd) return HS_ERR_INVALID;
  *ngroups = (uint32_t)ar->view.num_groups;
  *nxcd = (uint32_t)hs_host::xcd_groups(num_wgs);
  return HS_OK;
}

int hs_persistent_launch(hs_arena *ar, const void *kernel, int kind, int num_wgs, int wg_size,
                         size_t lds_bytes, void **args, void *stream) {
  if (!ar || !kernel || !args || kind < 0 || kind >= HS_BAR_KIND_COUNT || num_wgs <= 0 ||
      num_wgs > ar->view.max_blocks || wg_size <= 0 || wg_size > 1024)
    return HS_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  int rc = hs_host::check_residency(kernel, num_wgs, wg_size, lds_bytes);
  if (rc) return rc;
  if ((rc = hs_host::prepare_barrier(ar, kind, s))) return rc;
  return hs_host::launch_persistent(kernel, args, kind, num_wgs, wg_size, lds_bytes, s);
}

// ---- device information ----
int hs_device_info_get(hs_device_info *info) {
  if (!info) return HS_ERR_INVALID;
  memset(info, 0, sizeof(*info));
  int dev = 0;
  HS_TRY_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HS_TRY_HIP(hipGetDeviceProperties(&p, dev));
  info->num_cus = p.multiProcessorCount;
  info->num_xcds = hs_host::device_xcds();
  info->max_threads_per_block = p.maxThreadsPerBlock;
  info->lds_per_cu = (int)p.maxSharedMemoryPerMultiProcessor;
  info->warp_size = p.warpSize;
  info->coop_launch = p.cooperativeLaunch;
  info->hbm_bytes = (long long)p.totalGlobalMem;
  strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
  return HS_OK;
}

int hs_probe_xcc_ids(int num_wgs, int wg_size, int *h_out) {
  if (num_wgs <= 0 || wg_size <= 0 || wg_size > 1024 || !h_out) return HS_ERR_INVALID;
  int *d = nullptr;
  HS_TRY_HIP(hipMalloc(&d, sizeof(int) * num_wgs));
  hipLaunchKernelGGL(k_probe_xcc, dim3(num_wgs), dim3(wg_size), 0, 0, d);
  hipError_t e = hipMemcpy(h_out, d, sizeof(int) * num_wgs, hipMemcpyDeviceToHost);
  if (hipFree(d) != hipSuccess && e == hipSuccess) e = hipErrorUnknown;
  return e == hipSuccess ? HS_OK : HS_ERR_HIP;
}

const char *hs_status_string(int s) {
  switch (s) {
    case HS_OK: return "ok";
    case HS_ERR_INVALID: return "invalid argument";
    case HS_ERR_HIP: return "HIP runtime error";
    case HS_ERR_RESIDENCY: return "grid exceeds co-resident workgroups";
    case HS_ERR_ABORTED: return "aborted by host";
    case HS_ERR_TIMEOUT: return "device wait exceeded spin limit";
    case HS_ERR_NOT_INIT: return "locks not initialised";
    case HS_ERR_NOMEM: return "out of memory";
    case HS_ERR_CHECK: return "invariant violated";
    default: return "unknown status";
  }
}

const char *hs_version(void) { return "libhs 0.1 gfx950"; }

}  // extern "C"
