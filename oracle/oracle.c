/*
 * oracle.c -- CPU restatement of the reference's algorithms and host checks.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker.  The product never links it.
 *
 * Pinning (each function is checked against something that is not itself):
 *   orc_rand64 / orc_gen_uniform_u32 : SplitMix64 published test vectors
 *       (seed 1234567 sequence, Vigna's splitmix64.c) in tests/test_oracle.py,
 *       plus the committed fixture tests/golden/gen_uniform.json.
 *   orc_reduce_*      : closed forms (all-ones -> N as in Reduce/test_reduce.cu:14-17,
 *                       i-ramp -> N(N-1)/2), numpy int64 sum, int32 wrap.
 *   orc_scan_*        : all-ones -> i / i+1 (scan/main.cu:36-41 input), numpy cumsum,
 *                       inclusive = exclusive + in.
 *   orc_hs_golden     : closed form -- every (line_words*(L+1))-word segment ends up
 *                       shifted min(L, shifts) lines; brute-force tiny cases; a
 *                       planted double-shift is detected.
 *   orc_bfs           : path (level=i), star, disconnected (INF), scipy
 *                       shortest_path(unweighted) on random graphs.
 *   orc_sssp          : unit weights == BFS levels, scipy dijkstra on random graphs.
 *   orc_pagerank_f64  : cycle / complete graph -> 1/n, star closed form,
 *                       dense numpy eigen-solve on random graphs.
 *   orc_pagerank_push_fixpoint_f64 : dense numpy linear solve of the same equation.
 *   orc_gr_*          : hand-written byte fixture tests/golden/tiny.gr, round trip.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Counter-based generator.  SplitMix64 (Steele, Lea, Flood 2014; constants as
 * published by Vigna).  Element i of stream `seed` is mix64(mix64(seed) + (i+1)*G)
 * with G the golden-ratio increment, i.e. the (i+1)-th output of a SplitMix64
 * sequence started at state mix64(seed).  No reference generator exists (the
 * reference feeds all-ones arrays, Reduce/test_reduce.cu:14, scan/main.cu:39).  */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t orc_mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_rand64(uint64_t seed, uint64_t index)
{
  uint64_t state = orc_mix64(seed) + (index + 1) * ORC_GOLDEN;
  return orc_mix64(state);
}

void orc_gen_uniform_u32(uint64_t seed, uint64_t n, uint32_t lo, uint32_t hi,
                         uint32_t *out)
{
  const uint64_t range = (uint64_t)hi - lo + 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t r = orc_rand64(seed, (uint64_t)i) >> 32;
    out[i] = lo + (uint32_t)((r * range) >> 32);
  }
}

/* ------------------------------------------------------------------------ */
/* Reduce: the reference sums int with int arithmetic (Reduce/reduce_gsrb.cu:303-322),
 * i.e. two's-complement wrap; correct_sum for all-ones is N (Reduce/test_reduce.cu:14-17). */
int32_t orc_reduce_i32(const int32_t *in, uint64_t n)
{
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) acc += (uint32_t)in[i];
  return (int32_t)acc;
}

int64_t orc_reduce_i64(const int32_t *in, uint64_t n)
{
  int64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) acc += in[i];
  return acc;
}

/* Scan: cpu_sum_scan, exclusive, unsigned wrap (scan/main.cu:11-21). */
void orc_scan_exclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out)
{
  uint32_t run = 0;
  for (uint64_t i = 0; i < n; ++i) {
    out[i] = run;
    run = run + in[i];
  }
}

void orc_scan_inclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out)
{
  uint32_t run = 0;
  for (uint64_t i = 0; i < n; ++i) {
    run = run + in[i];
    out[i] = run;
  }
}

/* ------------------------------------------------------------------------ */
/* Heterosync.  Names and ids follow Heterosync/main.cu:2481-2532; the first
 * match wins, so "spinSemEBO*" resolves to 12-15 and ids 32-35 are unreachable
 * by name, as in the reference. */
static const struct { const char *name; int id; } orc_prims[] = {
  {"atomicTreeBarrUniq", 0}, {"atomicTreeBarrUniqLocalExch", 1},
  {"lfTreeBarrUniq", 2}, {"lfTreeBarrUniqLocalExch", 3},
  {"spinMutex", 4}, {"spinMutexEBO", 5}, {"sleepMutex", 6}, {"faMutex", 7},
  {"spinSem1", 8}, {"spinSem2", 9}, {"spinSem10", 10}, {"spinSem120", 11},
  {"spinSemEBO1", 12}, {"spinSemEBO2", 13}, {"spinSemEBO10", 14}, {"spinSemEBO120", 15},
  {"PRspinSem1", 16}, {"PRspinSem2", 17}, {"PRspinSem10", 18}, {"PRspinSem120", 19},
  {"spinMutexUniq", 20}, {"spinMutexEBOUniq", 21}, {"sleepMutexUniq", 22}, {"faMutexUniq", 23},
  {"spinSemUniq1", 24}, {"spinSemUniq2", 25}, {"spinSemUniq10", 26}, {"spinSemUniq120", 27},
  {"spinSemEBOUniq1", 28}, {"spinSemEBOUniq2", 29}, {"spinSemEBOUniq10", 30}, {"spinSemEBOUniq120", 31},
  {"spinSemEBO1", 32}, {"spinSemEBO2", 33}, {"spinSemEBO10", 34}, {"spinSemEBO120", 35},
  {"atomicTreeBarrUniqSRB", 37}, {"atomicTreeBarrUniqLocalExchSRB", 38},
  {"atomicTreeBarrUniqNaiveSRB", 39}, {"atomicTreeBarrUniqLocalExchNaiveSRB", 40},
  {"atomicTreeBarrUniqNaiveAllSRB", 41}, {"atomicTreeBarrUniqLocalExchNaiveAllSRB", 42},
  {"atomicTreeBarrUniqHybridSRB", 43}, {"atomicTreeBarrUniqLocalExchHybridSRB", 44},
  {"atomicTreeBarrUniqCCG", 45}, {"atomicTreeBarrUniqLocalExchCCG", 46},
};

int orc_hs_prim_id(const char *name)
{
  for (size_t i = 0; i < sizeof(orc_prims) / sizeof(orc_prims[0]); ++i)
    if (strcmp(orc_prims[i].name, name) == 0) return orc_prims[i].id;
  return -1;
}

static int orc_ceil_div(int a, int b) { return (a + b - 1) / b; }

/* numLocsMult (Heterosync/main.cu:2558-2579) */
int orc_hs_num_locs_mult(int prim, int num_tbs, int num_sm)
{
  const int per_sm = orc_ceil_div(num_tbs, num_sm);
  if (prim < 4 || prim >= 37 || (prim >= 24 && prim <= 35)) return num_tbs;
  if (prim >= 4 && prim <= 7) return 1;
  if (prim >= 8 && prim <= 19) return per_sm;
  if ((prim >= 20 && prim <= 23) || prim == 36) return num_tbs < num_sm ? num_tbs : num_sm;
  return -1;
}

/* accessData_golden (Heterosync/main.cu:41-59) with the cache line width and the
 * half-warp width as a parameter (the reference fixes both at 16 words). */
static void orc_access_golden(float *g, int64_t cur, int64_t nlocs, int num_ldst,
                              int line_words)
{
  if (cur % ((int64_t)line_words * (num_ldst + 1)) >= line_words) {
    if (cur - line_words < 0 || cur >= nlocs) abort();
    g[cur] = g[cur - line_words];
  }
}

static int orc_is_local_exch(int prim)
{
  return prim == 1 || prim == 3 || prim == 38 || prim == 40 || prim == 42 ||
         prim == 44 || prim == 46;
}

/* Golden replay, Heterosync/main.cu:2988-3144.  `wg_size` generalises the
 * reference's NUM_THREADS_PER_BLOCK (32); line width stays 16 words. */
void orc_hs_golden(int prim, int num_ldst, int num_tbs, int wg_size, int num_sm,
                   int num_iters, int num_repeats, float *g, int64_t nlocs)
{
  const int line_words = 16;
  const int per_sm = orc_ceil_div(num_tbs, num_sm);
  const int64_t P = (int64_t)wg_size * (num_ldst + 1); /* numUniqLocsAccPerTB */
  for (int64_t i = 0; i < nlocs; ++i) g[i] = (float)i;
  for (int rep = 0; rep < num_repeats; ++rep) {
    for (int j = 0; j < num_iters; ++j) {
      if (prim < 4 || prim >= 37) {
        const int64_t n = (int64_t)num_tbs * P;
        for (int64_t i = n - 1; i >= 0; --i)
          orc_access_golden(g, (i + (int64_t)j * P) % n, nlocs, num_ldst, line_words);
        if (orc_is_local_exch(prim) && per_sm > 1) {
          for (int64_t i = n - 1; i >= 0; --i) {
            int64_t cur = (i + (int64_t)per_sm * P) % n;
            cur = (cur + (int64_t)j * P) % n;
            orc_access_golden(g, cur, nlocs, num_ldst, line_words);
          }
        }
      } else if (prim >= 4 && prim <= 7) {
        for (int b = 0; b < num_tbs; ++b)
          for (int64_t i = P - 1; i >= 0; --i)
            orc_access_golden(g, i, nlocs, num_ldst, line_words);
      } else if ((prim >= 8 && prim <= 19) || (prim >= 32 && prim <= 35)) {
        const int nsm = num_tbs < num_sm ? num_tbs : num_sm;
        for (int tb = 0; tb < num_tbs; ++tb) {
          const int sm = tb % nsm, per = tb / nsm;
          if (per != sm % per_sm) continue;
          for (int k = 0; k < per_sm; ++k)
            for (int64_t i = P - 1; i >= 0; --i)
              orc_access_golden(g, (i + (int64_t)k * P) % nlocs, nlocs, num_ldst, line_words);
        }
      } else if (prim >= 20 && prim <= 23) {
        const int64_t n = (int64_t)num_tbs * P;
        for (int64_t i = n - 1; i >= 0; --i)
          orc_access_golden(g, i % ((int64_t)num_sm * P), nlocs, num_ldst, line_words);
      } else {
        const int64_t n = (int64_t)num_tbs * P;
        for (int64_t i = n - 1; i >= 0; --i)
          orc_access_golden(g, i, nlocs, num_ldst, line_words);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* BFS: level-synchronous semantics of LONESTAR/apps/bfs/kernel_gsrb.cu
 * (init 299-311: level[src]=0, others INF=INT_MAX; visit 473-479:
 * if level[dst]==INF then level[dst]=LEVEL).  Textbook queue BFS gives the same
 * levels because levels are schedule-invariant. */
void orc_bfs(int32_t n, const uint32_t *row, const uint32_t *col, int32_t src,
             int32_t *level)
{
  for (int32_t v = 0; v < n; ++v) level[v] = INT32_MAX;
  if (src < 0 || src >= n) return;
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t head = 0, tail = 0;
  level[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    int32_t u = q[head++];
    for (uint32_t e = row[u]; e < row[u + 1]; ++e) {
      int32_t v = (int32_t)col[e];
      if (level[v] == INT32_MAX) {
        level[v] = level[u] + 1;
        q[tail++] = v;
      }
    }
  }
  free(q);
}

/* SSSP: the near-far relaxation of LONESTAR/apps/sssp/kernel_gsrb.cu:415
 * (atomicMin(dist[dst], dist[n]+w)) converges to the unique shortest-path
 * distances; computed here by binary-heap Dijkstra. */
typedef struct { int64_t d; int32_t v; } orc_heap_item;

static void orc_heap_push(orc_heap_item *h, int64_t *sz, orc_heap_item it)
{
  int64_t i = (*sz)++;
  h[i] = it;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    orc_heap_item t = h[p]; h[p] = h[i]; h[i] = t;
    i = p;
  }
}

static orc_heap_item orc_heap_pop(orc_heap_item *h, int64_t *sz)
{
  orc_heap_item top = h[0];
  h[0] = h[--(*sz)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < *sz && h[l].d < h[m].d) m = l;
    if (r < *sz && h[r].d < h[m].d) m = r;
    if (m == i) break;
    orc_heap_item t = h[m]; h[m] = h[i]; h[i] = t;
    i = m;
  }
  return top;
}

void orc_sssp(int32_t n, const uint32_t *row, const uint32_t *col,
              const uint32_t *w, int32_t src, int32_t *dist)
{
  int64_t *d = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t v = 0; v < n; ++v) d[v] = INT64_MAX;
  if (src >= 0 && src < n) {
    int64_t cap = (int64_t)row[n] + 1, sz = 0;
    orc_heap_item *h = (orc_heap_item *)malloc(sizeof(orc_heap_item) * (size_t)cap);
    d[src] = 0;
    orc_heap_item s = {0, src};
    orc_heap_push(h, &sz, s);
    while (sz > 0) {
      orc_heap_item it = orc_heap_pop(h, &sz);
      if (it.d != d[it.v]) continue;
      for (uint32_t e = row[it.v]; e < row[it.v + 1]; ++e) {
        int32_t v = (int32_t)col[e];
        int64_t nd = it.d + (int64_t)w[e];
        if (nd < d[v]) {
          d[v] = nd;
          orc_heap_item ni = {nd, v};
          orc_heap_push(h, &sz, ni);
        }
      }
    }
    free(h);
  }
  for (int32_t v = 0; v < n; ++v)
    dist[v] = d[v] == INT64_MAX ? INT32_MAX : (d[v] >= INT32_MAX ? INT32_MAX - 1 : (int32_t)d[v]);
  free(d);
}

/* Normalized PageRank.  The equation (stated, since the reference's ranks are
 * un-normalized and drop dangling mass, LONESTAR/apps/pr/kernel_gsrb.cu:310-317):
 *   x'[v] = (1-a)/n + a * ( sum_{u->v} x[u]/deg(u) + D/n ),  D = sum_{deg(u)=0} x[u]
 * starting from x = 1/n, until ||x'-x||_1 < tol. Returns iterations. */
int orc_pagerank_f64(int32_t n, const uint32_t *row, const uint32_t *col,
                     double a, double tol, int max_iter, double *x)
{
  double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t v = 0; v < n; ++v) x[v] = 1.0 / n;
  int it = 0;
  for (; it < max_iter; ++it) {
    double dang = 0.0;
    for (int32_t u = 0; u < n; ++u)
      if (row[u + 1] == row[u]) dang += x[u];
    const double base = (1.0 - a) / n + a * dang / n;
    for (int32_t v = 0; v < n; ++v) y[v] = 0.0;
    for (int32_t u = 0; u < n; ++u) {
      const uint32_t deg = row[u + 1] - row[u];
      if (!deg) continue;
      const double c = x[u] / deg;
      for (uint32_t e = row[u]; e < row[u + 1]; ++e) y[col[e]] += c;
    }
    double l1 = 0.0;
    for (int32_t v = 0; v < n; ++v) {
      const double nv = base + a * y[v];
      l1 += nv > x[v] ? nv - x[v] : x[v] - nv;
      x[v] = nv;
    }
    if (l1 < tol) { ++it; break; }
  }
  free(y);
  return it;
}

/* Fixed point of the reference's residual push: p = (1-a) + a*sum p[u]/deg(u)
 * (init_1 p=1-a at kernel_gsrb.cu:310; pushes res*a/deg at 382-508). */
int orc_pagerank_push_fixpoint_f64(int32_t n, const uint32_t *row,
                                   const uint32_t *col, double a, double tol,
                                   int max_iter, double *p)
{
  double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t v = 0; v < n; ++v) p[v] = 1.0 - a;
  int it = 0;
  for (; it < max_iter; ++it) {
    for (int32_t v = 0; v < n; ++v) y[v] = 0.0;
    for (int32_t u = 0; u < n; ++u) {
      const uint32_t deg = row[u + 1] - row[u];
      if (!deg) continue;
      const double c = p[u] / deg;
      for (uint32_t e = row[u]; e < row[u + 1]; ++e) y[col[e]] += c;
    }
    double mx = 0.0;
    for (int32_t v = 0; v < n; ++v) {
      const double nv = (1.0 - a) + a * y[v];
      const double d = nv > p[v] ? nv - p[v] : p[v] - nv;
      if (d > mx) mx = d;
      p[v] = nv;
    }
    if (mx < tol) { ++it; break; }
  }
  free(y);
  return it;
}

/* ------------------------------------------------------------------------ */
/* .gr v1 (LONESTAR/rt/src/csr_graph.cu:187-199): u64 version=1, u64 edge type
 * size, u64 nnodes, u64 nedges, u64 row_end[nnodes], u32 dst[nedges], 4-byte pad
 * when nedges is odd, u32 data[nedges] when the edge type size is non-zero.
 * Little-endian host assumed (x86_64). */
int orc_gr_write(const char *path, uint32_t nnodes, uint64_t nedges,
                 const uint32_t *row, const uint32_t *col, const uint32_t *w)
{
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  uint64_t hdr[4] = {1, w ? 4u : 0u, nnodes, nedges};
  fwrite(hdr, 8, 4, f);
  for (uint32_t v = 0; v < nnodes; ++v) {
    uint64_t end = row[v + 1];
    fwrite(&end, 8, 1, f);
  }
  fwrite(col, 4, nedges, f);
  if (w) {
    if (nedges % 2) { uint32_t z = 0; fwrite(&z, 4, 1, f); }
    fwrite(w, 4, nedges, f);
  }
  return fclose(f) != 0;
}

int orc_gr_read_header(const char *path, uint64_t *nnodes, uint64_t *nedges,
                       uint64_t *edge_ty_size)
{
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  uint64_t hdr[4];
  size_t got = fread(hdr, 8, 4, f);
  fclose(f);
  if (got != 4 || hdr[0] != 1) return 2;
  *edge_ty_size = hdr[1];
  *nnodes = hdr[2];
  *nedges = hdr[3];
  return 0;
}

int orc_gr_read(const char *path, uint32_t *row, uint32_t *col, uint32_t *w)
{
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  uint64_t hdr[4];
  if (fread(hdr, 8, 4, f) != 4 || hdr[0] != 1) { fclose(f); return 2; }
  const uint64_t n = hdr[2], m = hdr[3];
  row[0] = 0;
  for (uint64_t v = 0; v < n; ++v) {
    uint64_t end;
    if (fread(&end, 8, 1, f) != 1) { fclose(f); return 3; }
    row[v + 1] = (uint32_t)end;
  }
  if (fread(col, 4, m, f) != m) { fclose(f); return 4; }
  if (hdr[1] && w) {
    if (m % 2) { uint32_t z; if (fread(&z, 4, 1, f) != 1) { fclose(f); return 5; } }
    if (fread(w, 4, m, f) != m) { fclose(f); return 6; }
  }
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU baselines: the same workloads on the host cores (OpenMP). */
int64_t orc_par_reduce_i64(const int32_t *in, uint64_t n, int *threads_used)
{
  int64_t acc = 0;
  int nt = 1;
#pragma omp parallel reduction(+ : acc)
  {
#ifdef _OPENMP
#pragma omp single
    nt = omp_get_num_threads();
#endif
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) acc += in[i];
  }
  if (threads_used) *threads_used = nt;
  return acc;
}

/* Host barrier episode: every OpenMP thread copies 16 floats within its own slab
 * (the GPU episode's 16 stores per thread), then joins a sense-reversing
 * spin barrier (one atomic counter + sense word, the G-CPUSRB shape).
 * Returns nanoseconds per episode; *threads_used receives the thread count. */
double orc_cpu_barrier_ns(int episodes, int *threads_used)
{
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  float *slab = (float *)calloc((size_t)nt * 64, sizeof(float));
  volatile int count = 0;
  volatile int sense = 0;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int local = 0;
    float *s = slab + (size_t)t * 64;
    for (int e = 0; e < episodes; ++e) {
      for (int k = 15; k >= 0; --k) s[k + 1] = s[k] + 1.0f;
      local = !local;
      if (__atomic_add_fetch(&count, 1, __ATOMIC_ACQ_REL) == nt) {
        __atomic_store_n(&count, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&sense, local, __ATOMIC_RELEASE);
      } else {
        while (__atomic_load_n(&sense, __ATOMIC_ACQUIRE) != local) __builtin_ia32_pause();
      }
    }
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(slab);
  if (threads_used) *threads_used = nt;
  const double ns = (t1.tv_sec - t0.tv_sec) * 1e9 + (t1.tv_nsec - t0.tv_nsec);
  return episodes > 0 ? ns / episodes : 0.0;
}

/* two-pass blocked scan: per-thread block sums, serial carry scan, block scan */
void orc_par_scan_inclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out,
                                int *threads_used)
{
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  uint32_t *sums = (uint32_t *)calloc((size_t)nt + 1, sizeof(uint32_t));
#pragma omp parallel num_threads(nt)
  {
    int t = 0, T = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    T = omp_get_num_threads();
#endif
    const uint64_t chunk = (n + T - 1) / T;
    const uint64_t b = (uint64_t)t * chunk < n ? (uint64_t)t * chunk : n;
    const uint64_t e = b + chunk < n ? b + chunk : n;
    uint32_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += in[i];
    sums[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int k = 1; k <= T; ++k) sums[k] += sums[k - 1];
    uint32_t run = sums[t];
    for (uint64_t i = b; i < e; ++i) { run += in[i]; out[i] = run; }
  }
  free(sums);
  if (threads_used) *threads_used = nt;
}

/* Direction-optimizing BFS (Beamer, Asanovic, Patterson, "Direction-Optimizing
 * Breadth-First Search", SC 2012): top-down from the frontier queue while the
 * frontier's out-edges are few (m_f <= m_u / 14), bottom-up over the unvisited
 * vertices' in-edges (reverse CSR) while they are many, back to top-down when the
 * frontier shrinks below n / 24 vertices.  Levels equal orc_bfs's (levels do
 * not depend on the traversal order).  Sources are spread over the OpenMP
 * threads, one source per thread at a time; levels[s*n + v]. */
static void orc_dobfs_one(int32_t n, const uint32_t *row, const uint32_t *col,
                          const uint32_t *rrow, const uint32_t *rcol, int32_t src,
                          int32_t *lv, int32_t *q, int32_t *nq, uint64_t *bits)
{
  for (int32_t v = 0; v < n; ++v) lv[v] = INT32_MAX;
  if (src < 0 || src >= n) return;
  const int64_t m = (int64_t)row[n];
  const int64_t words = ((int64_t)n + 63) / 64;
  lv[src] = 0;
  q[0] = src;
  int64_t nf = 1;
  int64_t mf = (int64_t)(row[src + 1] - row[src]);
  int64_t mu = m - mf;
  int bottom_up = 0;
  for (int32_t depth = 0; nf > 0; ++depth) {
    if (!bottom_up && mf * 14 > mu) bottom_up = 1;
    else if (bottom_up && nf * 24 < n) bottom_up = 0;
    int64_t nn = 0, nm = 0;
    if (!bottom_up) {
      for (int64_t i = 0; i < nf; ++i) {
        const int32_t u = q[i];
        for (uint32_t e = row[u]; e < row[u + 1]; ++e) {
          const int32_t v = (int32_t)col[e];
          if (lv[v] == INT32_MAX) {
            lv[v] = depth + 1;
            nq[nn++] = v;
            nm += (int64_t)(row[v + 1] - row[v]);
          }
        }
      }
    } else {
      memset(bits, 0, sizeof(uint64_t) * (size_t)words);
      for (int64_t i = 0; i < nf; ++i) bits[q[i] >> 6] |= 1ull << (q[i] & 63);
      for (int32_t v = 0; v < n; ++v) {
        if (lv[v] != INT32_MAX) continue;
        for (uint32_t e = rrow[v]; e < rrow[v + 1]; ++e) {
          const uint32_t u = rcol[e];
          if (bits[u >> 6] >> (u & 63) & 1) {
            lv[v] = depth + 1;
            nq[nn++] = v;
            nm += (int64_t)(row[v + 1] - row[v]);
            break;
          }
        }
      }
    }
    int32_t *t = q; q = nq; nq = t;
    nf = nn;
    mf = nm;
    mu -= nm;
  }
}

void orc_cpu_msbfs(int32_t n, const uint32_t *row, const uint32_t *col,
                   const uint32_t *rrow, const uint32_t *rcol, const int32_t *srcs,
                   int nsrc, int32_t *levels, int *threads_used)
{
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
#pragma omp parallel num_threads(nt)
  {
    const size_t nn = (size_t)(n > 0 ? n : 1);
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * nn);
    int32_t *nq = (int32_t *)malloc(sizeof(int32_t) * nn);
    uint64_t *bits = (uint64_t *)malloc(sizeof(uint64_t) * ((nn + 63) / 64));
#pragma omp for schedule(dynamic, 1)
    for (int s = 0; s < nsrc; ++s)
      orc_dobfs_one(n, row, col, rrow, rcol, srcs[s], levels + (size_t)s * nn, q, nq, bits);
    free(q);
    free(nq);
    free(bits);
  }
  if (threads_used) *threads_used = nt;
}

/* Binary-heap Dijkstra per source (orc_sssp), sources spread over the OpenMP
 * threads; dist[s*n + v]. */
void orc_cpu_mssp(int32_t n, const uint32_t *row, const uint32_t *col,
                  const uint32_t *w, const int32_t *srcs, int nsrc, int32_t *dist,
                  int *threads_used)
{
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int s = 0; s < nsrc; ++s)
    orc_sssp(n, row, col, w, srcs[s], dist + (size_t)s * (size_t)(n > 0 ? n : 1));
  if (threads_used) *threads_used = nt;
}

/* Pull power iteration of the normalized PageRank equation of orc_pagerank_f64
 * in FP32 storage (ranks and per-vertex contributions x[u]/deg(u)), FP64 sums
 * per vertex and for the L1 change; in-edges from the reverse CSR.  OpenMP over
 * vertices.  Returns the iterations run. */
int orc_cpu_pagerank_pull_f32(int32_t n, const uint32_t *row, const uint32_t *rrow,
                              const uint32_t *rcol, double a, double tol, int max_iter,
                              float *x, int *threads_used)
{
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  const size_t nn = (size_t)(n > 0 ? n : 1);
  float *c = (float *)malloc(sizeof(float) * nn);
  float *ybuf = (float *)malloc(sizeof(float) * nn);
  float *const out = x;
  float *y = ybuf;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int32_t v = 0; v < n; ++v) x[v] = (float)(1.0 / n);
  int it = 0;
  for (; it < max_iter; ++it) {
    double dang = 0.0;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : dang)
    for (int32_t u = 0; u < n; ++u) {
      const uint32_t deg = row[u + 1] - row[u];
      c[u] = deg ? (float)((double)x[u] / deg) : 0.0f;
      if (!deg) dang += x[u];
    }
    const double base = (1.0 - a) / n + a * dang / n;
    double l1 = 0.0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096) reduction(+ : l1)
    for (int32_t v = 0; v < n; ++v) {
      double s = 0.0;
      for (uint32_t e = rrow[v]; e < rrow[v + 1]; ++e) s += c[rcol[e]];
      const float nv = (float)(base + a * s);
      l1 += fabs((double)nv - (double)x[v]);
      y[v] = nv;
    }
    float *t = x; x = y; y = t;
    if (l1 < tol) { ++it; break; }
  }
  if (x != out) memcpy(out, x, sizeof(float) * nn);
  free(c);
  free(ybuf);
  if (threads_used) *threads_used = nt;
  return it;
}
