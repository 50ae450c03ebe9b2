// cc_labelprop.hip -- connected components by label propagation, written only
// against libhs's public headers and C ABI, as an application of one's own is:
// the frontier queues and wave expansion of hs_worklist.hpp, the grid barrier
// of hs_device.hpp, and hs_arena_device_view / hs_barrier_fanin /
// hs_persistent_launch of hs.h.  It plays the part of a further LonestarGPU app
// on the Worklist2 / PipeContextT runtime (LONESTAR/rt/include/worklist.h:49-585,
// LONESTAR/rt/include/pipe.h:29-123); libhs itself does not use this file.
//
// On a symmetric graph (every edge in both directions) the result is, for each
// vertex, the smallest vertex id of its component.  Level k takes the vertices
// whose label dropped in level k - 1 (all vertices at level 0), pushes each
// one's label to its neighbours with an atomic minimum, and appends every
// neighbour whose label dropped to the next frontier, once per level (a stamp
// per vertex).  Labels only decrease, so a stale read of a neighbour's label is
// corrected by the vertex that lowered it, which is in the next frontier.
#include <hip/hip_runtime.h>

#include "hs.h"
#include "hs_device.hpp"
#include "hs_worklist.hpp"

namespace {

constexpr int kWg = 256;
constexpr int kSteps = 4;          // 64-edge steps in flight per wave
constexpr uint32_t kStride = 64;   // words between frontier counters (own lines)

struct CcArgs {
  uint32_t n;
  const uint32_t *row, *col;
  uint32_t *label;   // [n] output
  uint32_t *stamp;   // [n] the last level + 1 that queued the vertex
  uint32_t *items;   // [2][n] frontier queues
  uint32_t *ctr;     // [6 * kStride] counts and grab counters
  uint32_t *levels;  // [1] (may be null): levels run
  uint32_t ng, nx;
};

__device__ __forceinline__ uint32_t grid_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t grid_threads() { return gridDim.x * blockDim.x; }

template <int KIND>
__global__ void __launch_bounds__(kWg) __attribute__((amdgpu_waves_per_eu(8))) k_cc(hs::ArenaView a, CcArgs p) {
  hs::GridBarrier gb;
  gb.init(a, KIND, p.ng, p.nx);
  const hs::Frontier fr{p.items, p.ctr, p.ctr + 3 * kStride, p.n, kStride};
  for (uint32_t v = grid_tid(); v < p.n; v += grid_threads()) {
    p.label[v] = v;
    p.stamp[v] = 0;
    p.items[v] = v;  // level 0 input: every vertex
  }
  if (grid_tid() == 0) {
    for (uint32_t i = 0; i < 3; ++i) {
      hs::st_relaxed(fr.count_at(i), i == 0 ? p.n : 0u);
      hs::st_relaxed(fr.grab_at(i), 0u);
    }
  }
  if (gb.sync<KIND>()) return;
  uint32_t k = 0;
  for (;; ++k) {
    const uint32_t nin = fr.in_count(k);
    if (nin == 0) break;
    const uint32_t *qin = fr.in(k);
    uint32_t *qout = fr.out(k);
    uint32_t *cout = fr.out_count(k);
    if (grid_tid() == 0) fr.clear_for_next(k);
    __shared__ uint32_t s_base;
    for (;;) {
      if (threadIdx.x == 0) s_base = fr.grab(k, (uint32_t)kWg);
      __syncthreads();
      const uint32_t base = s_base;
      __syncthreads();
      if (base >= nin) break;
      const uint32_t i = base + threadIdx.x;
      const bool valid = i < nin;
      const uint32_t u = valid ? qin[i] : 0u;
      uint32_t b = 0, deg = 0, lu = 0xFFFFFFFFu;
      if (valid) {
        b = p.row[u];
        deg = p.row[u + 1] - b;
        lu = hs::ld_relaxed(p.label + u);
      }
      hs::wave_expand<kSteps>(b, deg, [&](const uint32_t (&e)[kSteps], const bool (&ev)[kSteps],
                                          const uint32_t (&own)[kSteps]) {
        uint32_t w[kSteps];
        bool want[kSteps];
#pragma unroll
        for (int x = 0; x < kSteps; ++x) {
          const uint32_t l = __shfl(lu, (int)own[x]);
          w[x] = ev[x] ? p.col[e[x]] : 0u;
          want[x] = false;
          if (ev[x]) {
            const uint32_t old = __hip_atomic_fetch_min(p.label + w[x], l, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (l < old)
              want[x] = __hip_atomic_exchange(p.stamp + w[x], k + 1, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT) != k + 1;
          }
        }
        hs::wave_push_batch<kSteps>(qout, cout, want, w);
      });
    }
    if (gb.sync<KIND>()) return;
  }
  if (grid_tid() == 0 && p.levels) p.levels[0] = k;
}

template <int KIND>
const void *kernel_of() { return (const void *)k_cc<KIND>; }

const void *cc_kernel(int kind) {
  switch (kind) {
    case HS_BAR_FLAT: return kernel_of<HS_BAR_FLAT>();
    case HS_BAR_TWO_LEVEL: return kernel_of<HS_BAR_TWO_LEVEL>();
    case HS_BAR_XCD: return kernel_of<HS_BAR_XCD>();
    case HS_BAR_XCD_FLAT_RELEASE: return kernel_of<HS_BAR_XCD_FLAT_RELEASE>();
    case HS_BAR_COOP: return kernel_of<HS_BAR_COOP>();
    default: return nullptr;
  }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

// scratch bytes of hs_example_cc for a graph of n vertices
int hs_example_cc_work_size(uint32_t n, uint64_t *bytes) {
  if (!bytes) return HS_ERR_INVALID;
  *bytes = align256(sizeof(uint32_t) * (size_t)n) + align256(sizeof(uint32_t) * 2 * (size_t)n) +
           align256(sizeof(uint32_t) * 6 * kStride);
  return HS_OK;
}

// labels[v] = the smallest vertex id of v's component (symmetric CSR row/col);
// d_levels (device, may be NULL) receives the number of levels.  num_wgs 0 = one
// workgroup per CU.  Barrier kinds: flat, two-level, XCD, XCD flat release, coop.
int hs_example_cc(hs_arena *arena, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
                  uint32_t *d_labels, void *d_work, int barrier_kind, int num_wgs, uint32_t *d_levels,
                  void *stream) {
  const void *k = cc_kernel(barrier_kind);
  if (!arena || !k || n == 0 || !d_row || (m && !d_col) || !d_labels || !d_work || m > 0xFFFFFFFFull)
    return HS_ERR_INVALID;
  if (num_wgs <= 0) {
    hs_device_info info;
    int rc = hs_device_info_get(&info);
    if (rc) return rc;
    num_wgs = info.num_cus;
  }
  CcArgs p{};
  p.n = n;
  p.row = d_row;
  p.col = d_col;
  p.label = d_labels;
  char *w = (char *)d_work;
  p.stamp = (uint32_t *)w;
  w += align256(sizeof(uint32_t) * (size_t)n);
  p.items = (uint32_t *)w;
  w += align256(sizeof(uint32_t) * 2 * (size_t)n);
  p.ctr = (uint32_t *)w;
  p.levels = d_levels;
  int rc = hs_barrier_fanin(arena, num_wgs, &p.ng, &p.nx);
  if (rc) return rc;
  hs::ArenaView v{};
  if ((rc = hs_arena_device_view(arena, &v, sizeof(v)))) return rc;
  void *args[] = {(void *)&v, (void *)&p};
  return hs_persistent_launch(arena, k, barrier_kind, num_wgs, kWg, 0, args, stream);
}

}  // extern "C"
