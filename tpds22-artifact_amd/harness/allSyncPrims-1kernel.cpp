// allSyncPrims-1kernel -- the Heterosync microbenchmark harness on libhs.
//
// Same command line and output lines as Heterosync/main.cu:2372-3180:
//   ./allSyncPrims-1kernel <syncPrim> <numLdSt> <numTBs> <numCSIters> [wgSize]
// The optional fifth argument sets the workgroup size (the reference fixes 32
// threads per block, Heterosync/main.cu:8; a wave64 workgroup is 64 or more).
// The kernel runs NUM_REPEATS = 10 times on the same storage, then the host
// replays the data movement (golden) and compares at 1e-5.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hip_check.hpp"
#include "hs.h"

namespace {

const int kRepeats = 10;
const int kLineWords = 16;

const struct { const char *name; int id; const char *label; } kPrims[] = {
    {"atomicTreeBarrUniq", 0, "atomic_tree_barrier"},
    {"atomicTreeBarrUniqLocalExch", 1, "atomic_tree_barrier_localExch"},
    {"lfTreeBarrUniq", 2, "fbs_tree_barrier"},
    {"lfTreeBarrUniqLocalExch", 3, "fbs_tree_barrier_localExch"},
    {"spinMutex", 4, "spin_lock_mutex"}, {"spinMutexEBO", 5, "ebo_mutex"},
    {"sleepMutex", 6, "sleeping_mutex"}, {"faMutex", 7, "fetchadd_mutex"},
    {"spinSem1", 8, "spin_lock_sem"}, {"spinSem2", 9, "spin_lock_sem"},
    {"spinSem10", 10, "spin_lock_sem"}, {"spinSem120", 11, "spin_lock_sem"},
    {"spinSemEBO1", 12, "ebo_sem"}, {"spinSemEBO2", 13, "ebo_sem"},
    {"spinSemEBO10", 14, "ebo_sem"}, {"spinSemEBO120", 15, "ebo_sem"},
    {"PRspinSem1", 16, "spin_lock_sem"}, {"PRspinSem2", 17, "spin_lock_sem"},
    {"PRspinSem10", 18, "spin_lock_sem"}, {"PRspinSem120", 19, "spin_lock_sem"},
    {"spinMutexUniq", 20, "spin_lock_mutex_uniq"}, {"spinMutexEBOUniq", 21, "ebo_mutex_uniq"},
    {"sleepMutexUniq", 22, "sleeping_mutex_uniq"}, {"faMutexUniq", 23, "fetchadd_mutex_uniq"},
    {"spinSemUniq1", 24, "spin_lock_sem_uniq"}, {"spinSemUniq2", 25, "spin_lock_sem_uniq"},
    {"spinSemUniq10", 26, "spin_lock_sem_uniq"}, {"spinSemUniq120", 27, "spin_lock_sem_uniq"},
    {"spinSemEBOUniq1", 28, "ebo_sem_uniq"}, {"spinSemEBOUniq2", 29, "ebo_sem_uniq"},
    {"spinSemEBOUniq10", 30, "ebo_sem_uniq"}, {"spinSemEBOUniq120", 31, "ebo_sem_uniq"},
    {"atomicTreeBarrUniqSRB", 37, "atomic_tree_barrier_SRB"},
    {"atomicTreeBarrUniqLocalExchSRB", 38, "atomic_tree_barrier_localExch_SRB"},
    {"atomicTreeBarrUniqNaiveSRB", 39, "atomic_tree_barrier_NaiveSRB"},
    {"atomicTreeBarrUniqLocalExchNaiveSRB", 40, "atomic_tree_barrier_localExch_NaiveSRB"},
    {"atomicTreeBarrUniqNaiveAllSRB", 41, "atomic_tree_barrier_NaiveAllSRB"},
    {"atomicTreeBarrUniqLocalExchNaiveAllSRB", 42, "atomic_tree_barrier_localExch_NaiveAllSRB"},
    {"atomicTreeBarrUniqHybridSRB", 43, "atomic_tree_barrier_hybridSRB"},
    {"atomicTreeBarrUniqLocalExchHybridSRB", 44, "atomic_tree_barrier_localExch_hybridSRB"},
    {"atomicTreeBarrUniqCCG", 45, "atomic_tree_barrier_ccg"},
    {"atomicTreeBarrUniqLocalExchCCG", 46, "atomic_tree_barrier_localExch_ccg"},
};
const int kSemCaps[4] = {1, 2, 10, 120};

int locs_mult(int prim, int tbs, int nsm) {
  const int per = (tbs + nsm - 1) / nsm;
  if (prim < 4 || prim >= 37 || (prim >= 24 && prim <= 35)) return tbs;
  if (prim <= 7) return 1;
  if (prim <= 19) return per;
  return tbs < nsm ? tbs : nsm;
}

void access_golden(std::vector<float> &g, long cur, int L) {
  if (cur % ((long)kLineWords * (L + 1)) >= kLineWords) g[cur] = g[cur - kLineWords];
}

bool local_exch(int p) { return p == 1 || p == 3 || p == 38 || p == 40 || p == 42 || p == 44 || p == 46; }

// host replay of the kernels' data movement (Heterosync/main.cu:2988-3144)
void golden(int prim, int L, int tbs, int wg, int nsm, int iters, std::vector<float> &g) {
  const long P = (long)wg * (L + 1);
  const long nl = (long)g.size();
  const int per = (tbs + nsm - 1) / nsm;
  for (long i = 0; i < nl; ++i) g[i] = (float)i;
  for (int r = 0; r < kRepeats; ++r)
    for (int j = 0; j < iters; ++j) {
      if (prim < 4 || prim >= 37) {
        const long n = tbs * P;
        for (long i = n - 1; i >= 0; --i) access_golden(g, (i + j * P) % n, L);
        if (local_exch(prim) && per > 1)
          for (long i = n - 1; i >= 0; --i) access_golden(g, (((i + per * P) % n) + j * P) % n, L);
      } else if (prim <= 7) {
        for (int b = 0; b < tbs; ++b)
          for (long i = P - 1; i >= 0; --i) access_golden(g, i, L);
      } else if (prim <= 19) {
        const int ns = tbs < nsm ? tbs : nsm;
        for (int tb = 0; tb < tbs; ++tb) {
          const int sm = tb % ns, pid = tb / ns;
          if (pid != sm % per) continue;
          for (int k = 0; k < per; ++k)
            for (long i = P - 1; i >= 0; --i) access_golden(g, (i + k * P) % nl, L);
        }
      } else if (prim <= 23) {
        for (long i = tbs * P - 1; i >= 0; --i) access_golden(g, i % ((long)nsm * P), L);
      } else {
        for (long i = tbs * P - 1; i >= 0; --i) access_golden(g, i, L);
      }
    }
}

#define CHECK(x)                                                            \
  do {                                                                      \
    int _rc = (x);                                                          \
    if (_rc) {                                                              \
      fprintf(stderr, "ERROR: %s: %s\n", #x, hs_status_string(_rc));        \
      exit(-1);                                                             \
    }                                                                       \
  } while (0)

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "./allSyncPrims-1kernel <syncPrim> <numLdSt> <numTBs> <numCSIters> [wgSize]\n");
    fprintf(stderr, "\t<syncPrim>: one of\n");
    for (const auto &p : kPrims) fprintf(stderr, "\t\t%s\n", p.name);
    exit(-1);
  }
  hs_device_info info;
  CHECK(hs_device_info_get(&info));
  const int NUM_SM = info.num_cus;
  const int wg = argc == 6 ? atoi(argv[5]) : 64;
  const int maxTBPerSM = info.max_threads_per_block / wg;
  const int MAX_BLOCKS = maxTBPerSM * NUM_SM;
  fprintf(stdout, "GPU: %s, # XCD: %d\n", info.arch, info.num_xcds);
  fprintf(stdout, "# SM: %d, Max Thrs/TB: %d, Max TB/SM: %d, Max # TB: %d\n", NUM_SM,
          info.max_threads_per_block, maxTBPerSM, MAX_BLOCKS);
  const char *name = argv[1];
  const int L = atoi(argv[2]), tbs = atoi(argv[3]), iters = atoi(argv[4]);
  if (tbs <= 0 || tbs > MAX_BLOCKS || L < 0 || iters < 0 || wg <= 0 || wg % 64) {
    fprintf(stderr, "ERROR: bad numLdSt / numTBs / numCSIters / wgSize\n");
    exit(-1);
  }
  int prim = -1;
  const char *label = "";
  for (const auto &p : kPrims)
    if (!strcmp(p.name, name)) { prim = p.id; label = p.label; break; }
  if (prim < 0) {
    fprintf(stderr, "ERROR: Unknown synchronization primitive: %s\n", name);
    exit(-1);
  }
  // cudaLocksInit(MAX_BLOCKS, 8*NUM_SM, 24*NUM_SM, pageAlign, NUM_SM) (main.cu:2542)
  CHECK(hsLocksInit(MAX_BLOCKS, 8 * NUM_SM, 24 * NUM_SM, 0, NUM_SM));
  hs_arena *arena = hsLocksArena();
  const long nlocs = (long)locs_mult(prim, tbs, NUM_SM) * wg * (L + 1);
  fprintf(stdout, "# TBs: %d, # Ld/St: %d, # Locs Mult: %d, # Uniq Locs/TB: %d, # Storage Locs: %ld\n",
          tbs, L, locs_mult(prim, tbs, NUM_SM), wg * (L + 1), nlocs);
  if ((prim >= 8 && prim <= 19) || (prim >= 24 && prim <= 31))
    fprintf(stdout, "%s_%03d_%03d\n", label, kSemCaps[(prim - 8) % 4], iters);
  else
    fprintf(stdout, "%s_%03d\n", label, iters);
  std::vector<float> host(nlocs);
  for (long i = 0; i < nlocs; ++i) host[i] = (float)i;
  float *d = nullptr;
  uint32_t *viol = nullptr;
  if (hipMalloc(&d, sizeof(float) * nlocs) != hipSuccess ||
      hipMalloc(&viol, sizeof(uint32_t) * 64 * (NUM_SM + 1)) != hipSuccess) {
    fprintf(stderr, "ERROR: hipMalloc\n");
    exit(-1);
  }
  HIP_CHECK(hipMemcpy(d, host.data(), sizeof(float) * nlocs, hipMemcpyHostToDevice));
  hipEvent_t st, en;
  HIP_CHECK(hipEventCreate(&st));
  HIP_CHECK(hipEventCreate(&en));
  float total = 0.f, ms = 0.f;
  uint32_t violations = 0;
  for (int r = 0; r < kRepeats; ++r) {
    HIP_CHECK(hipEventRecord(st, 0));
    CHECK(hs_sync_bench(arena, prim, d, tbs, wg, L, iters, 1, viol, nullptr));
    HIP_CHECK(hipEventRecord(en, 0));
    CHECK(hs_arena_sync(arena, nullptr));
    HIP_CHECK(hipEventSynchronize(en));
    HIP_CHECK(hipEventElapsedTime(&ms, st, en));
    total += ms;
    uint32_t v = 0;
    HIP_CHECK(hipMemcpy(&v, viol, sizeof(v), hipMemcpyDeviceToHost));
    violations += v;
  }
  printf("elapsed time: %f ms, average: %f ms\n", ms, total / kRepeats);
  HIP_CHECK(hipMemcpy(host.data(), d, sizeof(float) * nlocs, hipMemcpyDeviceToHost));
  std::vector<float> g(nlocs);
  golden(prim, L, tbs, wg, NUM_SM, iters, g);
  fprintf(stdout, "Comparing GPU results to golden results:\n");
  unsigned errors = 0;
  for (long i = 0; i < nlocs; ++i)
    if (std::fabs(host[i] - g[i]) > 1e-5f) {
      if (errors < 10) fprintf(stderr, "\tERROR: storage[%ld] = %f, golden[%ld] = %f\n", i, host[i], i, g[i]);
      ++errors;
    }
  if (violations) fprintf(stderr, "ERROR: %s: %u invariant violations\n", name, violations);
  HIP_CHECK(hipFree(d));
  HIP_CHECK(hipFree(viol));
  hsLocksDestroy();
  if (errors || violations) {
    fprintf(stderr, "ERROR: %s has %u output errors\n", name, errors);
    exit(-1);
  }
  fprintf(stdout, "PASSED!\n");
  return 0;
}
