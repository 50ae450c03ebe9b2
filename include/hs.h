/*
 * hs.h -- C ABI of libhs, the MI355X (gfx950) device-wide synchronization
 * primitives and the persistent kernels that run on them.
 *
 * Conventions
 *   - every entry point returns an int status (hs_status); no exceptions cross
 *     the ABI;
 *   - buffers named d_* are caller-owned device pointers; h_* are host pointers;
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *   - an arena owns the lock/barrier state and is owned by its handle.
 *
 * The reference exposes a source-level CUDA API (.cu files #included into each
 * benchmark, Heterosync/main.cu:15-18) and no library; the entry points below
 * replace the host functions cited beside each one.
 */
#ifndef HS_H
#define HS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hs_status {
  HS_OK = 0,
  HS_ERR_INVALID = 1,    /* bad argument (also cudaErrorInitializationError in the reference) */
  HS_ERR_HIP = 2,        /* a HIP runtime call failed */
  HS_ERR_RESIDENCY = 3,  /* grid larger than what can be co-resident: refused */
  HS_ERR_ABORTED = 4,    /* the host raised the arena's abort flag */
  HS_ERR_TIMEOUT = 5,    /* a device spin exceeded the arena's spin limit */
  HS_ERR_NOT_INIT = 6,   /* singleton API used before hsLocksInit */
  HS_ERR_NOMEM = 7,
  HS_ERR_CHECK = 8       /* an in-kernel invariant counter is non-zero */
} hs_status;

/* barrier kinds (device-wide) */
typedef enum hs_barrier_kind {
  HS_BAR_FLAT = 0,       /* one counter + generation word: G-CPUSRB (Reduce/reduce_gcpusrb.cu, naive=true) */
  HS_BAR_TWO_LEVEL = 1,  /* per-group counter, then device counter: G-SRB (Reduce/reduce_gsrb.cu:135-269) */
  HS_BAR_XCD = 2,        /* per-XCD counter + per-XCD release word, then device counter */
  HS_BAR_XCD_FLAT_RELEASE = 3, /* per-XCD arrival counters, all WGs poll the device word */
  HS_BAR_TREE = 4,       /* HeteroSync atomic tree barrier (Heterosync/cudaLocksBarrierAtomic.cu:6-170) */
  HS_BAR_LF = 5,         /* lock-free flag tree barrier (Heterosync/cudaLocksBarrierFast.cu:10-330) */
  HS_BAR_COOP = 6,       /* cooperative-groups grid sync, hipLaunchCooperativeKernel */
  HS_BAR_NAIVE_ALL = 7,  /* every thread joins one SRB (Heterosync/cudaLocksBarrierAtomicSRB.cu:242-253) */
  HS_BAR_KIND_COUNT = 8
} hs_barrier_kind;

typedef enum hs_mutex_kind {
  HS_MUTEX_SPIN = 0,  /* Heterosync/cudaLocksMutexSpin.cu */
  HS_MUTEX_EBO = 1,   /* Heterosync/cudaLocksMutexEBO.cu */
  HS_MUTEX_FA = 2,    /* Heterosync/cudaLocksMutexFA.cu (ticket) */
  HS_MUTEX_SLEEP = 3  /* Heterosync/cudaLocksMutexSleep.cu (ring buffer) */
} hs_mutex_kind;

typedef enum hs_sem_kind {
  HS_SEM_SPIN = 0,          /* Heterosync/cudaLocksSemaphoreSpin.cu:36-409 */
  HS_SEM_EBO = 1,           /* Heterosync/cudaLocksSemaphoreEBO.cu:4-455 */
  HS_SEM_SPIN_PRIORITY = 2, /* Heterosync/cudaLocksSemaphoreSpin.cu:411-837 */
  HS_SEM_EBO_PRIORITY = 3,  /* Heterosync/cudaLocksSemaphoreEBO.cu:456-930 */
  HS_SEM_XCD_QUEUE = 4      /* per-XCD priority ticket queues (no reference counterpart) */
} hs_sem_kind;

typedef struct hs_arena hs_arena;

/* Replaces cudaLockData_t's sizing arguments (Heterosync/cudaLocks.h:4-16). */
typedef struct hs_arena_desc {
  int max_blocks;      /* maxBlocksPerKernel */
  int num_mutexes;     /* numMutexes */
  int num_semaphores;  /* numSemaphores */
  int num_groups;      /* NUM_SM: number of per-group ("per-SM") slots per lock */
  int page_align;      /* pageAlign: round every region to 4 KiB */
  uint32_t spin_limit; /* polls before a device wait gives up; 0 = default */
} hs_arena_desc;

/* ---- arena: replaces cudaLocksInit / cudaLocksDestroy (Heterosync/cudaLocks.cu:3-102) ---- */
int hs_arena_create(const hs_arena_desc *desc, hs_arena **out);
int hs_arena_destroy(hs_arena *arena);
/* resets every barrier, mutex and semaphore word to its initial value */
int hs_arena_reset(hs_arena *arena, void *stream);
/* raise / clear the abort flag polled by every device wait.  A wait looks at
   the flag after every 2048 polls, so waits that end sooner never pay for it. */
int hs_arena_set_abort(hs_arena *arena, int value);
/* enqueue on `stream` a device wait until *d_word == value (agent-scope load),
   bounded by the arena's spin limit and abort flag like every other wait; the
   outcome is reported by hs_arena_sync.  No reference counterpart: the stream
   analogue of the reference's spin on a flag (spinOnMyOutFlag,
   Heterosync/cudaLocksBarrierFast.cu:148-161). */
int hs_wait_value_u32(hs_arena *arena, const uint32_t *d_word, uint32_t value, void *stream);
/* status word written by the device (HS_OK, HS_ERR_TIMEOUT, HS_ERR_ABORTED) */
int hs_arena_device_status(hs_arena *arena, int *status_out);
/* launchers only enqueue work; this waits for `stream` and reports the device
   status of the arena's waits (HS_OK, HS_ERR_TIMEOUT, HS_ERR_ABORTED).  After a
   failure the arena is reset. */
int hs_arena_sync(hs_arena *arena, void *stream);
/* copies the first `count` per-workgroup episode slots to the host: with the
   barrier microbenchmark's check on, slot b holds the last episode workgroup b
   published before arriving.  Read after the device work ends and before
   hs_arena_sync, which clears them after a failed wait: a slot still at 0 names
   a workgroup that never arrived.  No reference counterpart (diagnostic). */
int hs_arena_episode_slots(hs_arena *arena, uint32_t *h_out, int count);
/* copies the arena's device view (hs::ArenaView of include/hs_device.hpp) for
   kernels written against the device API; view_bytes must be its size */
int hs_arena_device_view(hs_arena *arena, void *view_out, size_t view_bytes);

/* ---- handles: replace cudaMutexCreate{Spin,EBO,FA,Sleep} (e.g. Heterosync/cudaLocksMutexSpin.cu:4-9)
        and cudaSemaphoreCreate{Spin,EBO} (Heterosync/cudaLocksSemaphoreSpin.cu:6-29,
        Heterosync/cudaLocksSemaphoreEBO.cu:4-28) ---- */
int hs_mutex_create(hs_arena *arena, int kind, unsigned mutex_number, unsigned *handle);
int hs_sem_create(hs_arena *arena, int kind, unsigned sem_number, unsigned count, unsigned *handle);

/* ---- singleton API with the reference's names and argument meaning ---- */
int hsLocksInit(int maxBlocksPerKernel, int numMutexes, int numSemaphores, int pageAlign, int NUM_SM);
int hsLocksDestroy(void);
int hsMutexCreateSpin(unsigned *handle, int mutexNumber);
int hsMutexCreateEBO(unsigned *handle, int mutexNumber);
int hsMutexCreateFA(unsigned *handle, int mutexNumber);
int hsMutexCreateSleep(unsigned *handle, int mutexNumber);
int hsSemaphoreCreateSpin(unsigned *handle, int semaphoreNumber, unsigned count, int NUM_SM);
int hsSemaphoreCreateEBO(unsigned *handle, int semaphoreNumber, unsigned count, int NUM_SM);
hs_arena *hsLocksArena(void);

/* ---- device / launch information ---- */
typedef struct hs_device_info {
  int num_cus;
  int num_xcds;
  int max_threads_per_block;
  int lds_per_cu;
  int warp_size;
  int coop_launch;
  long long hbm_bytes;
  char arch[64];
} hs_device_info;
int hs_device_info_get(hs_device_info *info);
/* how many workgroups of `wg_size` threads of the given barrier-bench kernel fit at once */
int hs_max_resident_wgs(int barrier_kind, int wg_size, int *out);
/* physical XCD id of every workgroup of a grid (probe of WG->XCD dispatch) */
int hs_probe_xcc_ids(int num_wgs, int wg_size, int *h_xcc_out);
/* number of groups of HS_BAR_XCD / HS_BAR_XCD_FLAT_RELEASE (workgroup b joins
   group b % groups) for later launches in this process; 0 (default) = one group
   per XCD, doubled until no group has more than 255 members.  Any value is correct; a group whose members span several XCDs takes
   the barrier's write-back-request path.  No reference counterpart. */
int hs_set_xcd_groups(int groups);

/* ---- persistent kernels of one's own: the launch a LonestarGPU app writes by hand
   (grid capped at residency, LONESTAR/apps/bfs/kernel_gsrb.cu:653-655; barrier
   state allocated and cleared, LONESTAR/apps/bfs/kernel_gsrb.cu:672-690) for a kernel written
   against hs_device.hpp / hs_worklist.hpp.
   hs_barrier_fanin: the ngroups / nxcd arguments of hs::GridBarrier::init for a
   grid of num_wgs workgroups on this arena.
   hs_persistent_launch: refuses a grid that is not co-resident (HS_ERR_RESIDENCY),
   clears the barrier state `barrier_kind` needs, and launches `kernel` (a
   __global__ function of the caller's own code object; args as for
   hipLaunchKernel) on `stream`, cooperatively for HS_BAR_COOP.  The residency
   check trusts the runtime's occupancy figure, which assumes more SGPRs per wave
   than gfx950 grants: give such kernels __attribute__((amdgpu_waves_per_eu(8)))
   as libhs's own have (then their waves fit the slots the figure counts).
   Results and status as for libhs's launches: hs_arena_sync. */
int hs_barrier_fanin(hs_arena *arena, int num_wgs, uint32_t *ngroups, uint32_t *nxcd);
int hs_persistent_launch(hs_arena *arena, const void *kernel, int barrier_kind, int num_wgs,
                         int wg_size, size_t lds_bytes, void **args, void *stream);

/* ---- Heterosync microbenchmark (replaces main()/invoke* of Heterosync/main.cu:1500-2370) ----
   prim: the reference's syncPrim id (Heterosync/main.cu:2481-2532).
   d_storage: numLocsMult*wg_size*(num_ldst+1) floats (Heterosync/main.cu:2558-2590).
   One call = one kernel launch of `iterations` critical-section iterations.
   d_violations (may be NULL): per-episode invariant counters (barriers: early passes;
   mutex: overlapping owners; semaphore: capacity / writer exclusivity). */
int hs_sync_bench(hs_arena *arena, int prim, float *d_storage, int num_tbs, int wg_size,
                  int num_ldst, int iterations, int check, uint32_t *d_violations,
                  void *stream);
/* explicit barrier microbench: any hs_barrier_kind at any grid size */
int hs_barrier_bench(hs_arena *arena, int kind, float *d_storage, int num_wgs, int wg_size,
                     int num_ldst, int episodes, int check, uint32_t *d_violations,
                     void *stream);
/* semaphore microbench with reader:writer mix (writers = wg % ratio == 0).
   d_violations (may be NULL; >= 128 u32, zeroed by the call): [0] capacity /
   writer-exclusivity violations, [1] the longest wait of one acquisition in 10 ns
   ticks (saturating), [2..3] the sum of all acquisitions' waits in ticks (u64). */
int hs_sem_bench(hs_arena *arena, int kind, unsigned sem, float *d_storage, int num_wgs,
                 int wg_size, int num_ldst, int iterations, int writer_every,
                 uint32_t *d_violations, void *stream);

/* ---- persistent int32 reduce (replaces reduce(), Reduce/reduce_gsrb.cu:340-404) ----
   d_partials: scratch of >= num_wgs int32 (num_wgs <= 0: library default).
   Result (two's-complement wrap, like int arithmetic) written to d_out[0]. */
int hs_reduce_i32(hs_arena *arena, const int32_t *d_in, uint64_t n, int32_t *d_partials,
                  int32_t *d_out, int barrier_kind, int num_wgs, void *stream);
int hs_reduce_partials_size(int num_wgs, uint64_t n, uint64_t *elems_out);
/* the same reduce with in-kernel clock stamps (diagnostic, no reference
   counterpart; barrier kinds flat, two-level, XCD, coop): d_stamps (device,
   >= 4 x num_wgs u64) receives per workgroup b, at [4b], [4b+1], [4b+2], the
   device's 100 MHz clock at its start, when its loads are summed, and when it
   leaves (after the fold for the last arriver) -- the launch split into
   dispatch spread, streaming and fold. */
int hs_reduce_i32_stamped(hs_arena *arena, const int32_t *d_in, uint64_t n, int32_t *d_partials,
                          int32_t *d_out, int barrier_kind, int num_wgs, uint64_t *d_stamps,
                          void *stream);

/* ---- persistent u32/int32 scan (replaces sum_scan_blelloch, scan/scan_gsrb.cu:687-804) ----
   inclusive != 0 -> inclusive scan, else exclusive (the reference's). */
int hs_scan_u32(hs_arena *arena, const uint32_t *d_in, uint32_t *d_out, uint64_t n,
                int inclusive, uint32_t *d_scratch, int barrier_kind, int num_wgs,
                void *stream);
int hs_scan_scratch_size(int num_wgs, uint64_t n, uint64_t *elems_out);
/* hs_scan_u32 without d_scratch, for callers that cannot allocate device memory
   (code under stream capture, serving loops, harnesses that forbid hipMalloc).
   No reference counterpart.  The partials live in a static device array of the
   library (32 KiB per device, held until the process ends), so the call makes no
   allocation of its own; its address is looked up once per device by
   hs_arena_create.  Limits: num_wgs <= 4096 (the default grid is 2 per CU),
   kind != HS_BAR_COOP, the arena's device current.  Calls on one device share
   that array: calls that overlap on one GPU (two streams) race; calls in one
   stream are ordered. */
int hs_scan_u32_noalloc(hs_arena *arena, const uint32_t *d_in, uint32_t *d_out, uint64_t n,
                        int inclusive, int barrier_kind, int num_wgs, void *stream);
/* bytes of input per round of hs_scan_u32 in this process (default 64 MiB: a
   round stays resident in the Infinity Cache between its sum and scan passes) */
int hs_scan_set_round_bytes(uint64_t bytes);
int hs_scan_get_round_bytes(uint64_t *bytes_out);

/* ---- batched scans / reduces: many uneven segments in one persistent launch ----
   No reference counterpart (the reference runs one array per binary).  Segment s is
   [offsets[s], offsets[s+1]) of d_in; offsets are given on the device (d_offsets)
   and the host (h_offsets, used to size the launch).  Workgroups take 4096-element
   tiles in order from one counter (work stealing); the scan passes carries by a
   decoupled look-back.  d_work: hs_batch_work_size bytes.  nseg <= 65536. */
int hs_batch_work_size(const uint64_t *h_offsets, uint32_t nseg, uint64_t *bytes_out);
int hs_scan_batch_u32(const uint32_t *d_in, uint32_t *d_out, const uint64_t *d_offsets,
                      const uint64_t *h_offsets, uint32_t nseg, int inclusive, void *d_work,
                      int num_wgs, void *stream);
int hs_reduce_batch_i32(const int32_t *d_in, int32_t *d_sums, const uint64_t *d_offsets,
                        const uint64_t *h_offsets, uint32_t nseg, void *d_work, int num_wgs,
                        void *stream);
/* HS_OK, or HS_ERR_TIMEOUT when a look-back wait of the last launch on d_work gave up */
int hs_batch_status(const void *d_work, void *stream);

/* ---- device serving loop (no reference counterpart) ----
   One persistent kernel on a stream of its own takes requests from a ring in pinned
   host memory, admits each when its working-memory cost (mem_units) fits the budget
   (budget_units < 2^32-1: the device API's EBO priority semaphore, an admission a
   weighted reader of mem_units, completions posting with priority; an admission
   wait past the spin limit stops the loop), splits it into tiles that all
   workgroups share (continuous batching) and writes a completion record.
   type: 0 reduce_i32 (in: int32[n], out: int32[1]); 1 scan_u32 inclusive (in, out:
   uint32[n], n <= 2^28); 2 bfs (in: row, aux0: col, aux1: 2n u32 scratch, out:
   int32[n] levels, arg: source, n <= 2^28); 3 pagerank (in: reverse row, aux0:
   reverse col, aux2: forward row, aux1: 2n float scratch, out: float[n], arg:
   iterations, n <= 2^24).  All pointers are device pointers passed as integers.
   num_wgs 0 starts one workgroup per CU.  Completion status 0 is success, 5 a
   bounded wait that gave up or a full token queue.
   While a loop runs, nothing in the process may wait for the whole device
   (hipDeviceSynchronize, torch.cuda.synchronize, hipFree of device memory): such a
   wait includes the loop itself, which ends only at hs_serve_stop.  A loop of the
   co-resident maximum of workgroups (2 per CU) holds every CU's registers: other
   kernels launched meanwhile wait until it stops, so prepare inputs and buffers
   before hs_serve_start, or start it with fewer workgroups. */
typedef struct hs_server hs_server;
int hs_serve_start(int num_wgs, uint32_t ring_capacity, uint32_t budget_units, hs_server **out);
int hs_serve_submit(hs_server *server, uint32_t type, uint32_t mem_units, uint32_t arg, uint64_t in,
                    uint64_t out, uint64_t n, uint64_t aux0, uint64_t aux1, uint64_t aux2,
                    uint32_t *index_out);
/* PageRank request run to convergence: iterations stop once the L1 change of an
   iteration, sum |x_k - x_(k-1)|, is below tol (the stopping rule of the pull
   iteration that LONESTAR/apps/pr/kernel_gsrb.cu:291-292 parameterises with
   EPSILON 1e-6), or after max_iter iterations; tol 0 runs all max_iter.  The
   arguments are those of type 3 above. */
int hs_serve_submit_pagerank(hs_server *server, uint32_t mem_units, uint32_t max_iter, float tol,
                             uint64_t rrow, uint64_t rank, uint64_t n, uint64_t rcol, uint64_t scratch,
                             uint64_t row, uint32_t *index_out);
/* 1 when request `index` completed (status, device ticks at admission / end at
   100 MHz), 0 while pending */
int hs_serve_poll(hs_server *server, uint32_t index, uint32_t *status, uint64_t *t_admit,
                  uint64_t *t_end);
/* phases run by completed request `index` (pagerank: iterations + 1; bfs: 1 +
   the number of levels 0..max; reduce, scan: 1); HS_ERR_INVALID while it is
   pending or once its ring slot has been reused */
int hs_serve_phases(hs_server *server, uint32_t index, uint32_t *phases);
int hs_serve_stop(hs_server *server, int abort);

/* ---- hipGraph capture of repeated launches (the reference repeats launches in a
   host loop, Heterosync/main.cu:1507-1526).  Every libhs launcher can be captured:
   begin on a stream the caller created (not the null stream), enqueue launches,
   end to instantiate; replay re-runs the whole sequence as one submission.
   Launchers that size or check things on the host (residency, grids) do so at
   capture time; device status is still reported by hs_arena_sync. ---- */
typedef struct hs_graph hs_graph;
int hs_graph_capture_begin(void *stream);
int hs_graph_capture_end(void *stream, hs_graph **out);
int hs_graph_replay(hs_graph *graph, void *stream);
int hs_graph_destroy(hs_graph *graph);

/* ---- autotuner (the role of the LONESTAR/tuning scripts, which drive OpenTuner over
   the launch parameters of the reference's kernels) ----
   Times candidate configurations of one launch on the caller's input (each run
   `reps` times after a warm-up, HIP events on `stream`) and returns the fastest.
   Every configuration computes the same bits.  The launcher's default
   configuration is always timed first, so the result is never slower than the
   default as measured in the same call.  Candidates a device refuses (residency)
   or whose waits fail are skipped.  Allocates its own scratch (device memory). */
typedef struct {
  int algo;             /* scan: 0 = multi-phase barrier scan, 1 = single-pass look-back */
  int num_wgs;          /* 0 = the launcher's default grid */
  int barrier_kind;     /* hs_barrier_kind (barrier scan, reduce) */
  uint64_t round_bytes; /* barrier scan: hs_scan_set_round_bytes value */
  float ms;             /* per launch, the chosen configuration */
  float default_ms;     /* per launch, the default configuration */
  int tried;            /* configurations timed */
  int xcd_groups;       /* barrier fan-in: hs_set_xcd_groups value to apply (0 = default) */
} hs_tune_result;
/* hs_reduce_i32: workgroups per CU 1/2/4 x barrier kind XCD/two-level/flat, and
   the XCD barrier with 2x / 4x the groups (fewer members per group) */
int hs_tune_reduce_i32(hs_arena *arena, const int32_t *d_in, uint64_t n, int reps,
                       hs_tune_result *out, void *stream);
/* hs_barrier_bench with HS_BAR_XCD on the caller's workload (grid, workgroup
   size, stores per thread, episodes per launch): the barrier's fan-in, i.e. the
   XCD groups (the default, then 2x, 4x, 8x, 16x the XCD count while every group
   keeps at least 4 members); apply out->xcd_groups with hs_set_xcd_groups */
int hs_tune_barrier(hs_arena *arena, float *d_storage, int num_wgs, int wg_size, int num_ldst,
                    int episodes, int reps, hs_tune_result *out, void *stream);
/* hs_scan_u32 (grid 1/2/4 per CU x round 32/64/192 MiB, and 2x the XCD groups)
   and the look-back scan (2/4/8 per CU); the process's round size and XCD
   groups are left as they were */
int hs_tune_scan_u32(hs_arena *arena, const uint32_t *d_in, uint32_t *d_out, uint64_t n,
                     int inclusive, int reps, hs_tune_result *out, void *stream);
/* hs_scan_batch_u32: workgroups per CU 1/2/4/8 (the batch's carries pass by
   look-back, so there is no barrier to tune); out->num_wgs is the grid to pass */
int hs_tune_scan_batch_u32(const uint32_t *d_in, uint32_t *d_out, const uint64_t *d_offsets,
                           const uint64_t *h_offsets, uint32_t nseg, int inclusive, int reps,
                           hs_tune_result *out, void *stream);

/* ---- schedule MLP (8 -> 64 -> 64 -> 64 -> 8, ReLU, FP32), trained on the GPU ----
   No reference counterpart: the network scores the 8 actions of a multi-source
   BFS level inside hs_msbfs (schedule 3): action = direction (0 push / 1 pull)
   | expansion on half of the workgroups << 1 | long barrier poll backoff << 2.
   Samples are 8 features per row.  Two losses: labels (one action per sample,
   weights scale each sample's cross-entropy, normalised by wsum) and costs (the
   expected cost sum_a softmax(o)_a * cost[s][a] / norm, with cost [n][8] the
   measured time of every action). */
int hs_mlp_param_count(void);
int hs_mlp_scratch_floats(void);
int hs_mlp_init(float *d_w, uint64_t seed, void *stream);
int hs_mlp_forward(const float *d_w, const float *d_x, int n, float *d_logits, void *stream);
int hs_mlp_grad(const float *d_w, const float *d_x, const int *d_y, const float *d_wt, int n,
                float wsum, float *d_grad, float *d_loss, float *d_scratch, void *stream);
int hs_mlp_train_step(float *d_w, const float *d_x, const int *d_y, const float *d_wt, int n,
                      float wsum, float lr, int step, float *d_loss, float *d_scratch, void *stream);
/* number of actions (network outputs) */
int hs_mlp_actions(void);
int hs_mlp_grad_cost(const float *d_w, const float *d_x, const float *d_cost, int n, float norm,
                     float *d_grad, float *d_loss, float *d_scratch, void *stream);
int hs_mlp_train_step_cost(float *d_w, const float *d_x, const float *d_cost, int n, float norm, float lr,
                           int step, float *d_loss, float *d_scratch, void *stream);

/* ---- input generators (counter-based SplitMix64; identical on host and device) ---- */
int hs_gen_uniform_u32(uint32_t *d_out, uint64_t n, uint64_t seed, uint32_t lo, uint32_t hi,
                       void *stream);
int hs_gen_uniform_u32_host(uint32_t *h_out, uint64_t n, uint64_t seed, uint32_t lo, uint32_t hi);

/* ---- graphs: CSR uint32 (LONESTAR/rt/include/csr_graph.h:22-103) ---- */
/* R-MAT edge list -> CSR on the host: self loops dropped, duplicates merged,
   neighbours ascending.  Call once with row/col NULL to get sizes. */
int hs_rmat_csr_host(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                     uint32_t *h_row, uint32_t *h_col, uint64_t *nedges_out);
int hs_gen_weights_host(uint32_t *h_w, uint64_t nedges, uint64_t seed, uint32_t lo, uint32_t hi);
/* .gr v1 reader/writer (LONESTAR/rt/src/csr_graph.cu:152-237) */
int hs_gr_read_header(const char *path, uint64_t *nnodes, uint64_t *nedges, uint64_t *edge_ty_size);
int hs_gr_read(const char *path, uint32_t *h_row, uint32_t *h_col, uint32_t *h_w);
int hs_gr_write(const char *path, uint32_t nnodes, uint64_t nedges, const uint32_t *h_row,
                const uint32_t *h_col, const uint32_t *h_w);

/* reverse (transpose) CSR on the host */
int hs_csr_transpose_host(uint32_t nnodes, uint64_t nedges, const uint32_t *h_row,
                          const uint32_t *h_col, uint32_t *h_rrow, uint32_t *h_rcol);

const char *hs_status_string(int status);
const char *hs_version(void);

#ifdef __cplusplus
}
#endif
#endif
