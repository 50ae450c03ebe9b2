// tune.hip -- native autotuner for the persistent launches.
//
// The reference tunes its kernels offline with OpenTuner scripts over their
// launch parameters (LONESTAR/tuning/bfstuner.py:9-16, the block factor of
// LONESTAR/apps/bfs/kernel_gsrb.cu:716-719).  Here the search space is the one
// that matters on this device for a persistent launch: grid size (workgroups
// per CU), barrier kind and barrier fan-in (the number of groups of the XCD
// barrier: members per group), and, for the scan, the algorithm and its round
// size.  Each candidate is timed on the caller's input with HIP events; all of
// them compute the same bits, so the choice only affects time.  The launcher's
// default is always the first candidate, so the choice is never slower than the
// default as measured in the same call.
#include <hip/hip_runtime.h>

#include <vector>

#include "hs_graph.h"
#include "hs_internal.hpp"

namespace {

struct Cand {
  int algo, num_wgs, kind;
  uint64_t round_bytes;
  int groups;  // XCD barrier groups (hs_set_xcd_groups value; 0 = the default)
};

// the XCD-group override in force on entry, restored on exit
struct GroupsGuard {
  int saved;
  GroupsGuard() : saved(hs_host::xcd_groups_override()) {}
  ~GroupsGuard() { (void)hs_set_xcd_groups(saved); }
};

struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ok = false;
  Timer() { ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess; }
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// average ms per launch of `reps` launches after one warm-up, or -1 when the
// configuration is refused or one of its waits failed
template <class F>
float time_launches(hs_arena *a, F launch, int reps, hipStream_t s, Timer &t) {
  if (launch() != HS_OK) return -1.f;
  if (hs_arena_sync(a, s) != HS_OK) return -1.f;
  if (hipEventRecord(t.e0, s) != hipSuccess) return -1.f;
  for (int i = 0; i < reps; ++i)
    if (launch() != HS_OK) {
      (void)hs_arena_sync(a, s);
      return -1.f;
    }
  if (hipEventRecord(t.e1, s) != hipSuccess) return -1.f;
  if (hs_arena_sync(a, s) != HS_OK) return -1.f;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, t.e0, t.e1) != hipSuccess) return -1.f;
  return ms / (float)reps;
}

template <class F>
int pick(hs_arena *a, const std::vector<Cand> &cands, F run, int reps, hipStream_t s,
         hs_tune_result *out) {
  Timer t;
  if (!t.ok) return HS_ERR_HIP;
  float best = -1.f, def = -1.f;
  int tried = 0;
  Cand bc{};
  GroupsGuard guard;
  for (size_t i = 0; i < cands.size(); ++i) {
    const float ms = time_launches(a, [&] {
      const int g = hs_set_xcd_groups(cands[i].groups ? cands[i].groups : guard.saved);
      return g ? g : run(cands[i]);
    }, reps, s, t);
    if (ms < 0.f) continue;
    ++tried;
    if (i == 0) def = ms;
    if (best < 0.f || ms < best) { best = ms; bc = cands[i]; }
  }
  if (def < 0.f) return HS_ERR_HIP;  // the default itself failed
  out->algo = bc.algo;
  out->num_wgs = bc.num_wgs;
  out->barrier_kind = bc.kind;
  out->round_bytes = bc.round_bytes;
  out->xcd_groups = bc.groups;
  out->ms = best;
  out->default_ms = def;
  out->tried = tried;
  return HS_OK;
}

}  // namespace

extern "C" {

int hs_tune_reduce_i32(hs_arena *a, const int32_t *d_in, uint64_t n, int reps, hs_tune_result *out,
                       void *stream) {
  if (!a || !d_in || !out || n == 0 || reps < 1) return HS_ERR_INVALID;
  const int cus = hs_host::device_cus();
  const int x = hs_host::device_xcds();
  std::vector<Cand> cands{{0, 0, HS_BAR_XCD, 0, 0}};
  for (int per_cu : {1, 2, 4}) {
    for (int kind : {HS_BAR_XCD, HS_BAR_TWO_LEVEL, HS_BAR_FLAT}) cands.push_back({0, per_cu * cus, kind, 0, 0});
    for (int gm : {2, 4}) cands.push_back({0, per_cu * cus, HS_BAR_XCD, 0, gm * x});  // fan-in
  }
  uint64_t np = 0;
  int rc = hs_reduce_partials_size(4 * cus, n, &np);
  if (rc) return rc;
  int32_t *part = nullptr, *res = nullptr;
  HS_TRY_HIP(hipMalloc(&part, sizeof(int32_t) * (np + 64)));
  if (hipMalloc(&res, sizeof(int32_t) * 64) != hipSuccess) {
    (void)hipFree(part);
    return HS_ERR_NOMEM;
  }
  hipStream_t s = (hipStream_t)stream;
  rc = pick(a, cands, [&](const Cand &c) { return hs_reduce_i32(a, d_in, n, part, res, c.kind, c.num_wgs, s); },
            reps, s, out);
  const hipError_t f1 = hipFree(part), f2 = hipFree(res);
  if (rc == HS_OK && (f1 != hipSuccess || f2 != hipSuccess)) rc = HS_ERR_HIP;
  return rc;
}

int hs_tune_barrier(hs_arena *a, float *d_storage, int num_wgs, int wg_size, int num_ldst, int episodes, int reps,
                    hs_tune_result *out, void *stream) {
  if (!a || !d_storage || !out || num_wgs <= 0 || episodes <= 0 || reps < 1) return HS_ERR_INVALID;
  // fan-in of the XCD barrier: the default groups (one per XCD, doubled until
  // no group has more than 255 members), then 2x to 16x the XCDs' count while
  // groups keep at least 4 members
  const int x = hs_host::device_xcds();
  std::vector<Cand> cands{{0, num_wgs, HS_BAR_XCD, 0, 0}};
  for (int gm : {2, 4, 8, 16})
    if (gm * x <= num_wgs / 4 && gm * x <= hs::kMaxGroups) cands.push_back({0, num_wgs, HS_BAR_XCD, 0, gm * x});
  hipStream_t s = (hipStream_t)stream;
  return pick(a, cands, [&](const Cand &c) {
    return hs_barrier_bench(a, c.kind, d_storage, c.num_wgs, wg_size, num_ldst, episodes, 0, nullptr, s);
  }, reps, s, out);
}

int hs_tune_scan_u32(hs_arena *a, const uint32_t *d_in, uint32_t *d_out, uint64_t n, int inclusive,
                     int reps, hs_tune_result *out, void *stream) {
  if (!a || !d_in || !d_out || !out || n == 0 || reps < 1) return HS_ERR_INVALID;
  const int cus = hs_host::device_cus();
  uint64_t round0 = 0;
  hs_scan_get_round_bytes(&round0);
  const int x = hs_host::device_xcds();
  std::vector<Cand> cands{{0, 0, HS_BAR_XCD, round0, 0}};
  for (int per_cu : {1, 2, 4})
    for (uint64_t mb : {32ull, 64ull, 192ull}) cands.push_back({0, per_cu * cus, HS_BAR_XCD, mb << 20, 0});
  for (int per_cu : {1, 2, 4}) cands.push_back({0, per_cu * cus, HS_BAR_XCD, round0, 2 * x});  // fan-in
  for (int per_cu : {2, 4, 8}) cands.push_back({1, per_cu * cus, 0, 0, 0});
  uint64_t nscr = 0, nwork = 0;
  int rc = hs_scan_scratch_size(4 * cus, n, &nscr);
  if (rc) return rc;
  const uint64_t h_off[2] = {0, n};
  rc = hs_batch_work_size(h_off, 1, &nwork);
  if (rc) return rc;
  uint32_t *scr = nullptr;
  void *work = nullptr;
  uint64_t *d_off = nullptr;
  if (hipMalloc(&scr, sizeof(uint32_t) * (nscr + 64)) != hipSuccess ||
      hipMalloc(&work, nwork + 256) != hipSuccess || hipMalloc(&d_off, sizeof(h_off)) != hipSuccess) {
    (void)hipFree(scr);
    (void)hipFree(work);
    (void)hipFree(d_off);
    return HS_ERR_NOMEM;
  }
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(d_off, h_off, sizeof(h_off), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    rc = HS_ERR_HIP;
  } else {
    rc = pick(a, cands, [&](const Cand &c) {
      if (c.algo == 1) return hs_scan_batch_u32(d_in, d_out, d_off, h_off, 1, inclusive, work, c.num_wgs, s);
      const int r = hs_scan_set_round_bytes(c.round_bytes);
      return r ? r : hs_scan_u32(a, d_in, d_out, n, inclusive, scr, c.kind, c.num_wgs, s);
    }, reps, s, out);
  }
  hs_scan_set_round_bytes(round0);
  const hipError_t f1 = hipFree(scr), f2 = hipFree(work), f3 = hipFree(d_off);
  if (rc == HS_OK && (f1 != hipSuccess || f2 != hipSuccess || f3 != hipSuccess)) rc = HS_ERR_HIP;
  return rc;
}

}  // extern "C"

namespace {

// candidates of the persistent graph launches: the default first, then grid
// (workgroups per CU) x barrier (XCD with one group per XCD, XCD with twice the
// groups, two-level)
std::vector<Cand> graph_cands() {
  const int cus = hs_host::device_cus(), x = hs_host::device_xcds();
  std::vector<Cand> c{{0, 0, HS_BAR_XCD, 0, 0}};
  for (int per_cu : {1, 2, 4, 8}) {
    c.push_back({0, per_cu * cus, HS_BAR_XCD, 0, 0});
    c.push_back({0, per_cu * cus, HS_BAR_XCD, 0, 2 * x});
    c.push_back({0, per_cu * cus, HS_BAR_TWO_LEVEL, 0, 0});
  }
  return c;
}

// device scratch owned by one tuning call
struct Scratch {
  void *p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);  // nothing to do if the free fails
  }
  bool alloc(uint64_t bytes) { return hipMalloc(&p, bytes ? bytes : 8) == hipSuccess; }
};

}  // namespace

extern "C" {

int hs_tune_bfs(hs_arena *a, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col, int32_t src,
                int32_t *d_level, int reps, hs_tune_result *out, void *stream) {
  if (!a || !d_row || !d_col || !d_level || !out || n == 0 || reps < 1) return HS_ERR_INVALID;
  uint64_t nb = 0;
  int rc = hs_bfs_work_size(n, m, &nb);
  if (rc) return rc;
  Scratch w;
  if (!w.alloc(nb)) return HS_ERR_NOMEM;
  hipStream_t s = (hipStream_t)stream;
  return pick(a, graph_cands(), [&](const Cand &c) {
    return hs_bfs(a, n, m, d_row, d_col, src, d_level, w.p, c.kind, c.num_wgs, nullptr, s);
  }, reps, s, out);
}

int hs_tune_msbfs(hs_arena *a, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
                  const uint32_t *d_rrow, const uint32_t *d_rcol, const int32_t *d_srcs, int nsrc,
                  int32_t *d_levels, int schedule, const float *d_mlp, int reps, hs_tune_result *out,
                  void *stream) {
  if (!a || !d_row || !d_col || !d_rrow || !d_rcol || !d_srcs || !d_levels || !out || n == 0 || reps < 1 ||
      nsrc < 1 || nsrc > 64 || schedule == 4)
    return HS_ERR_INVALID;
  uint64_t nb = 0;
  int rc = hs_msbfs_work_size(n, m, nsrc, &nb);
  if (rc) return rc;
  Scratch w;
  if (!w.alloc(nb)) return HS_ERR_NOMEM;
  hipStream_t s = (hipStream_t)stream;
  return pick(a, graph_cands(), [&](const Cand &c) {
    return hs_msbfs(a, n, m, d_row, d_col, d_rrow, d_rcol, d_srcs, nsrc, d_levels, w.p, c.kind, c.num_wgs,
                    schedule, d_mlp, nullptr, nullptr, s);
  }, reps, s, out);
}

int hs_tune_mssp(hs_arena *a, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_col,
                 const uint32_t *d_w, const int32_t *d_srcs, int nsrc, int32_t delta, int32_t *d_dist, int reps,
                 hs_tune_result *out, void *stream) {
  if (!a || !d_row || !d_col || !d_w || !d_srcs || !d_dist || !out || n == 0 || reps < 1 || nsrc < 1 ||
      nsrc > 64)
    return HS_ERR_INVALID;
  uint64_t nb = 0;
  int rc = hs_mssp_work_size(n, m, nsrc, &nb);
  if (rc) return rc;
  Scratch w;
  if (!w.alloc(nb)) return HS_ERR_NOMEM;
  hipStream_t s = (hipStream_t)stream;
  return pick(a, graph_cands(), [&](const Cand &c) {
    return hs_mssp(a, n, m, d_row, d_col, d_w, d_srcs, nsrc, delta, d_dist, w.p, c.kind, c.num_wgs, nullptr, s);
  }, reps, s, out);
}

int hs_tune_pagerank_f32(hs_arena *a, uint32_t n, uint64_t m, const uint32_t *d_row, const uint32_t *d_rrow,
                         const uint32_t *d_rcol, float alpha, float tol_l1, int max_iter, float *d_rank, int reps,
                         hs_tune_result *out, void *stream) {
  if (!a || !d_row || !d_rrow || !d_rcol || !d_rank || !out || n == 0 || reps < 1) return HS_ERR_INVALID;
  uint64_t nb = 0;
  int rc = hs_pagerank_work_size(n, m, &nb);
  if (rc) return rc;
  Scratch w;
  if (!w.alloc(nb)) return HS_ERR_NOMEM;
  hipStream_t s = (hipStream_t)stream;
  return pick(a, graph_cands(), [&](const Cand &c) {
    return hs_pagerank_f32(a, n, m, d_row, d_rrow, d_rcol, alpha, tol_l1, max_iter, d_rank, w.p, c.kind,
                           c.num_wgs, nullptr, s);
  }, reps, s, out);
}

// the batched scan has no arena (its carries pass by look-back, not barriers):
// grid only; the status the waits report is checked with hs_batch_status
int hs_tune_scan_batch_u32(const uint32_t *d_in, uint32_t *d_out, const uint64_t *d_offsets,
                           const uint64_t *h_offsets, uint32_t nseg, int inclusive, int reps,
                           hs_tune_result *out, void *stream) {
  if (!d_in || !d_out || !d_offsets || !h_offsets || !out || nseg == 0 || reps < 1) return HS_ERR_INVALID;
  uint64_t nb = 0;
  int rc = hs_batch_work_size(h_offsets, nseg, &nb);
  if (rc) return rc;
  Scratch w;
  if (!w.alloc(nb)) return HS_ERR_NOMEM;
  hipStream_t s = (hipStream_t)stream;
  const int cus = hs_host::device_cus();
  std::vector<Cand> cands{{1, 0, 0, 0, 0}};
  for (int per_cu : {1, 2, 4, 8}) cands.push_back({1, per_cu * cus, 0, 0, 0});
  Timer t;
  if (!t.ok) return HS_ERR_HIP;
  float best = -1.f, def = -1.f;
  int tried = 0;
  Cand bc{};
  for (size_t i = 0; i < cands.size(); ++i) {
    auto launch = [&] { return hs_scan_batch_u32(d_in, d_out, d_offsets, h_offsets, nseg, inclusive, w.p,
                                                 cands[i].num_wgs, s); };
    if (launch() != HS_OK || hs_batch_status(w.p, s) != HS_OK) continue;
    if (hipEventRecord(t.e0, s) != hipSuccess) return HS_ERR_HIP;
    bool ok = true;
    for (int r = 0; r < reps && ok; ++r) ok = launch() == HS_OK;
    if (hipEventRecord(t.e1, s) != hipSuccess) return HS_ERR_HIP;
    if (!ok || hs_batch_status(w.p, s) != HS_OK) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.e0, t.e1) != hipSuccess) return HS_ERR_HIP;
    ms /= (float)reps;
    ++tried;
    if (i == 0) def = ms;
    if (best < 0.f || ms < best) { best = ms; bc = cands[i]; }
  }
  if (def < 0.f) return HS_ERR_HIP;
  out->algo = 1;
  out->num_wgs = bc.num_wgs;
  out->barrier_kind = 0;
  out->round_bytes = 0;
  out->xcd_groups = 0;
  out->ms = best;
  out->default_ms = def;
  out->tried = tried;
  return HS_OK;
}

}  // extern "C"
