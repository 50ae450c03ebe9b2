/*
 * hs_graph.h -- C ABI of libhs's persistent graph kernels (BFS, multi-source
 * BFS, SSSP, PageRank) on CSR uint32 graphs.
 *
 * Graph layout (LONESTAR/rt/include/csr_graph.h:22-103): row[nnodes+1] offsets,
 * col[nedges] destinations, optional w[nedges] weights, all device pointers.
 * Scratch (`d_work`) is caller-owned device memory of the size the matching
 * *_work_size function returns.  Launchers enqueue on `stream`; results and
 * device status are final after hs_arena_sync(arena, stream).
 */
#ifndef HS_GRAPH_H
#define HS_GRAPH_H

#include "hs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* single-source BFS: replaces gg_main + gg_main_pipe_1_gpu_gb
   (LONESTAR/apps/bfs/kernel_gsrb.cu:591-730).  d_level[v] = hops from src,
   INT32_MAX when unreached.  d_stats (device, >= 4 u32, may be NULL) receives
   {levels, pushes, 0, 0}. */
int hs_bfs_work_size(uint32_t nnodes, uint64_t nedges, uint64_t *bytes_out);
int hs_bfs(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
           const uint32_t *d_col, int32_t src, int32_t *d_level, void *d_work, int barrier_kind,
           int num_wgs, uint32_t *d_stats, void *stream);

/* the same BFS with the reference's load balancer on (bfs -l: the inspector, the
   prefix sum over the degrees of the frontier vertices of out-degree >=
   DEGREE_LIMIT and the kernel that gives every thread an equal share of their
   edges, LONESTAR/apps/bfs/kernel_gsrb.cu:343-398,569-579,
   LONESTAR/rt/include/thread_work.h:9-79): vertices of out-degree >= degree_limit
   have their edges split evenly over every wave of the grid, the others are
   expanded as in hs_bfs.  degree_limit 0 = the grid's thread count (the
   reference's DEGREE_LIMIT, LONESTAR/rt/include/ggcuda.h:19-22), else >= 256.
   Same levels as hs_bfs; same d_work size (hs_bfs_work_size). */
int hs_bfs_lb(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
              const uint32_t *d_col, int32_t src, int32_t *d_level, void *d_work, int barrier_kind,
              int num_wgs, int degree_limit, uint32_t *d_stats, void *stream);

/* multi-source BFS: up to 64 sources per launch as the bits of one 64-bit word
   per vertex (one wave64 lane per source).  d_levels is [nsrc][nnodes] int32.
   Each level is either a push over the frontier's out-edges or a pull over the
   unvisited vertices' in-edges (d_rrow/d_rcol: the transposed CSR); the choice
   never changes the result.
   A level's action is its direction (bit 0: 0 push / 1 pull), whether only half
   of the workgroups expand it (bit 1; the others wait at the barrier) and the
   poll backoff of the barrier that ends it (bit 2: long); none changes the result.
   schedule: 0 = the plain heuristic (direction by frontier out-edges vs unvisited
               in-edges, all workgroups, short backoff),
             1 = always push, 2 = always pull,
             3 = learned: d_mlp holds hs_mlp_param_count() network parameters and
               then three floats, the judge: margin_small, margin_large and the
               size threshold t (in log2(nnodes + 1)); the network's best action
               is taken only when its score leads the heuristic's action by at
               least the margin of the graph's band (margin_small when
               log2(nnodes + 1) <= t, else margin_large), else the heuristic's
               (+inf: always the heuristic),
             4 = per-level actions from h_forced (one byte per level, bits as above).
   d_trace (may be NULL): per-level records of HS_MSBFS_TRACE_WORDS u32 each:
   {action, frontier vertices, frontier out-edges (lo,hi), unvisited in-edges (lo,hi),
    new discoveries, start clock lo, start clock hi, end clock lo, end clock hi, 0...}.
   num_wgs <= 0: 4 workgroups per CU (the kernel's occupancy), at most the
   arena's max_blocks. */
#define HS_MSBFS_MAX_LEVELS 256
#define HS_MSBFS_TRACE_WORDS 16
int hs_msbfs_work_size(uint32_t nnodes, uint64_t nedges, int nsrc, uint64_t *bytes_out);
int hs_msbfs(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
             const uint32_t *d_col, const uint32_t *d_rrow, const uint32_t *d_rcol,
             const int32_t *d_srcs, int nsrc, int32_t *d_levels, void *d_work, int barrier_kind,
             int num_wgs, int schedule, const float *d_mlp, const uint8_t *h_forced,
             uint32_t *d_trace, void *stream);

/* SSSP with near/far buckets of width delta (delta <= 0 -> 10000): replaces
   gg_main (LONESTAR/apps/sssp/kernel_gsrb.cu:795-820).  Distances are the
   unique shortest-path lengths (INT32_MAX when unreached). */
int hs_sssp_work_size(uint32_t nnodes, uint64_t nedges, uint64_t *bytes_out);
int hs_sssp(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
            const uint32_t *d_col, const uint32_t *d_w, int32_t src, int32_t delta,
            int32_t *d_dist, void *d_work, int barrier_kind, int num_wgs, uint32_t *d_stats,
            void *stream);

/* multi-source SSSP: up to 64 sources in one persistent launch (one wave64 lane
   per source, 64-bit change masks per vertex, delta-stepping with one bound for
   all sources; delta = INT32_MAX is frontier Bellman-Ford).  The batched form of
   hs_sssp / gg_main (LONESTAR/apps/sssp/kernel_gsrb.cu:795-820): d_dist is
   [nsrc][nnodes] int32, each row equal to hs_sssp from that source.  d_stats
   (device, >= 4 u32, may be NULL) receives {relaxation rounds, bound steps,
   expanded (vertex, source) pairs mod 2^32, 0}. */
int hs_mssp_work_size(uint32_t nnodes, uint64_t nedges, int nsrc, uint64_t *bytes_out);
int hs_mssp(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
            const uint32_t *d_col, const uint32_t *d_w, const int32_t *d_srcs, int nsrc, int32_t delta,
            int32_t *d_dist, void *d_work, int barrier_kind, int num_wgs, uint32_t *d_stats,
            void *stream);

/* PageRank, FP32, normalized (sum 1), dangling mass spread uniformly, pull over
   the transposed CSR in a fixed order (deterministic); stops when the L1 change
   of an iteration is below tol_l1 or after max_iter iterations.  Replaces
   gg_main (LONESTAR/apps/pr/kernel_gsrb.cu:690-716).  d_stats receives
   {iterations, last L1 change as float bits}. */
int hs_pagerank_work_size(uint32_t nnodes, uint64_t nedges, uint64_t *bytes_out);
int hs_pagerank_f32(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
                    const uint32_t *d_rrow, const uint32_t *d_rcol, float alpha, float tol_l1,
                    int max_iter, float *d_rank, void *d_work, int barrier_kind, int num_wgs,
                    uint32_t *d_stats, void *stream);

/* autotuners of the graph launches (the role of LONESTAR/tuning/bfstuner.py,
   which searches the block factor of LONESTAR/apps/bfs/kernel_gsrb.cu:716-719):
   workgroups per CU 1/2/4/8 x barrier (XCD, XCD with 2x the groups, two-level),
   each timed `reps` times after a warm-up on the caller's graph; the default
   configuration is timed first and the choice is never slower than it in the
   same call.  Every configuration computes the same result (it is left in the
   output buffer).  The process's XCD groups are left as they were; apply
   out->xcd_groups with hs_set_xcd_groups.  Allocate their own scratch. */
int hs_tune_bfs(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
                const uint32_t *d_col, int32_t src, int32_t *d_level, int reps, hs_tune_result *out,
                void *stream);
int hs_tune_msbfs(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
                  const uint32_t *d_col, const uint32_t *d_rrow, const uint32_t *d_rcol,
                  const int32_t *d_srcs, int nsrc, int32_t *d_levels, int schedule, const float *d_mlp,
                  int reps, hs_tune_result *out, void *stream);
int hs_tune_mssp(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
                 const uint32_t *d_col, const uint32_t *d_w, const int32_t *d_srcs, int nsrc, int32_t delta,
                 int32_t *d_dist, int reps, hs_tune_result *out, void *stream);
int hs_tune_pagerank_f32(hs_arena *arena, uint32_t nnodes, uint64_t nedges, const uint32_t *d_row,
                         const uint32_t *d_rrow, const uint32_t *d_rcol, float alpha, float tol_l1,
                         int max_iter, float *d_rank, int reps, hs_tune_result *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
