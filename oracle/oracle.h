/*
 * oracle.h -- CPU restatement of the reference's host-side checks and
 * algorithms for the synchronization-primitive hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py use it, and there only as the checker / CPU baseline.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Pinning (what checks each function against something other than itself) is
 * listed in the header comment of oracle.c.
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based input generator (same definition as the product's) ---- */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_rand64(uint64_t seed, uint64_t index);
void orc_gen_uniform_u32(uint64_t seed, uint64_t n, uint32_t lo, uint32_t hi,
                         uint32_t *out);

/* ---- reduce (Reduce/test_reduce.cu:14-24, Reduce/reduce_gsrb.cu:289-337) ---- */
int32_t orc_reduce_i32(const int32_t *in, uint64_t n);
int64_t orc_reduce_i64(const int32_t *in, uint64_t n);

/* ---- scan (scan/main.cu:11-21) ---- */
void orc_scan_exclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out);
void orc_scan_inclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out);

/* ---- Heterosync microbenchmark golden replay (Heterosync/main.cu) ---- */
/* categories of the reference's syncPrim ids (Heterosync/main.cu:2481-2532) */
int orc_hs_prim_id(const char *name);
int orc_hs_num_locs_mult(int prim, int num_tbs, int num_sm);
void orc_hs_golden(int prim, int num_ldst, int num_tbs, int wg_size,
                   int num_sm, int num_iters, int num_repeats,
                   float *golden, int64_t num_storage_locs);

/* ---- graphs (LONESTAR) ---- */
/* BFS levels: level[src]=0, unreached = INT32_MAX
   (LONESTAR/apps/bfs/kernel_gsrb.cu:295-311, 473-479) */
void orc_bfs(int32_t nnodes, const uint32_t *row, const uint32_t *col,
             int32_t src, int32_t *level);
/* SSSP distances by binary-heap Dijkstra, unreached = INT32_MAX
   (LONESTAR/apps/sssp/kernel_gsrb.cu:297-309, 415) */
void orc_sssp(int32_t nnodes, const uint32_t *row, const uint32_t *col,
              const uint32_t *w, int32_t src, int32_t *dist);
/* normalized PageRank by FP64 power iteration; dangling mass spread uniformly */
int orc_pagerank_f64(int32_t nnodes, const uint32_t *row, const uint32_t *col,
                     double alpha, double tol_l1, int max_iter, double *rank);
/* the reference's un-normalized residual-push fixed point
   p = (1-a) + a * sum_{u->v} p[u]/deg(u)  (LONESTAR/apps/pr/kernel_gsrb.cu:298-336,
   382-508), solved by FP64 Jacobi iteration */
int orc_pagerank_push_fixpoint_f64(int32_t nnodes, const uint32_t *row,
                                   const uint32_t *col, double alpha,
                                   double tol, int max_iter, double *rank);

/* .gr v1 file I/O (LONESTAR/rt/src/csr_graph.cu:187-199) */
int orc_gr_write(const char *path, uint32_t nnodes, uint64_t nedges,
                 const uint32_t *row, const uint32_t *col, const uint32_t *w);
int orc_gr_read_header(const char *path, uint64_t *nnodes, uint64_t *nedges,
                       uint64_t *edge_ty_size);
int orc_gr_read(const char *path, uint32_t *row, uint32_t *col, uint32_t *w);

/* ---- CPU baselines (OpenMP; threads used returned) ---- */
int64_t orc_par_reduce_i64(const int32_t *in, uint64_t n, int *threads_used);
void orc_par_scan_inclusive_u32(const uint32_t *in, uint64_t n, uint32_t *out,
                                int *threads_used);
double orc_cpu_barrier_ns(int episodes, int *threads_used);
/* direction-optimizing BFS, one source per thread: levels[s*n + v] */
void orc_cpu_msbfs(int32_t nnodes, const uint32_t *row, const uint32_t *col,
                   const uint32_t *rrow, const uint32_t *rcol, const int32_t *srcs,
                   int nsrc, int32_t *levels, int *threads_used);
/* Dijkstra, one source per thread: dist[s*n + v] */
void orc_cpu_mssp(int32_t nnodes, const uint32_t *row, const uint32_t *col,
                  const uint32_t *w, const int32_t *srcs, int nsrc, int32_t *dist,
                  int *threads_used);
/* pull power iteration of orc_pagerank_f64's equation, FP32 ranks; returns iterations */
int orc_cpu_pagerank_pull_f32(int32_t nnodes, const uint32_t *row, const uint32_t *rrow,
                              const uint32_t *rcol, double alpha, double tol, int max_iter,
                              float *x, int *threads_used);

#ifdef __cplusplus
}
#endif
#endif
