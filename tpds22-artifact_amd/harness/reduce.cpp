// reduce -- Reduce/test_reduce.cu on libhs: ./reduce N threads_per_block [barrier [reps]]
// Input all ones, prints the kernel time, "Sum" and "Correct sum"
// (Reduce/test_reduce.cu:8-28, Reduce/reduce_gsrb.cu:391-396) and, unlike the
// reference, exits non-zero when they differ.  With reps > 1 the launch is
// captured once into a hipGraph (hs_graph_capture_*) and replayed reps times;
// the time printed is then the average per replay.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hip_check.hpp"
#include "hs.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s N threads_per_block [flat|two_level|xcd|coop [reps]]\n", argv[0]);
    return 2;
  }
  const long N = atol(argv[1]);
  const int kind = argc > 3 ? (!strcmp(argv[3], "flat") ? HS_BAR_FLAT
                               : !strcmp(argv[3], "two_level") ? HS_BAR_TWO_LEVEL
                               : !strcmp(argv[3], "coop") ? HS_BAR_COOP : HS_BAR_XCD)
                            : HS_BAR_XCD;
  (void)atoi(argv[2]);  // the persistent kernel always runs 256-thread workgroups
  const int reps = argc > 4 ? atoi(argv[4]) : 1;
  if (reps < 1) return 2;
  if (N < 0) return 2;
  hs_device_info info;
  if (hs_device_info_get(&info)) return 1;
  hs_arena_desc d = {8 * info.num_cus, 1, 1, info.num_cus, 0, 0};
  hs_arena *a = nullptr;
  if (hs_arena_create(&d, &a)) return 1;
  std::vector<int32_t> arr(N, 1);
  const int32_t correct = (int32_t)N;
  int32_t *din = nullptr, *dpart = nullptr, *dout = nullptr;
  uint64_t np = 0;
  hs_reduce_partials_size(0, N, &np);
  if (hipMalloc(&din, sizeof(int32_t) * (N ? N : 1)) || hipMalloc(&dpart, sizeof(int32_t) * np) ||
      hipMalloc(&dout, sizeof(int32_t)))
    return 1;
  HIP_CHECK(hipMemcpy(din, arr.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
  hipEvent_t s, e;
  HIP_CHECK(hipEventCreate(&s));
  HIP_CHECK(hipEventCreate(&e));
  int rc = HS_OK;
  if (reps == 1) {
    HIP_CHECK(hipEventRecord(s, 0));
    rc = hs_reduce_i32(a, din, N, dpart, dout, kind, 0, nullptr);
    HIP_CHECK(hipEventRecord(e, 0));
    if (!rc) rc = hs_arena_sync(a, nullptr);
  } else {
    hipStream_t st;
    HIP_CHECK(hipStreamCreate(&st));
    hs_graph *g = nullptr;
    rc = hs_reduce_i32(a, din, N, dpart, dout, kind, 0, st);  // warm-up (loads the code object)
    if (!rc) rc = hs_arena_sync(a, st);
    if (!rc) rc = hs_graph_capture_begin(st);
    if (!rc) {
      const int r1 = hs_reduce_i32(a, din, N, dpart, dout, kind, 0, st);
      const int r2 = hs_graph_capture_end(st, &g);
      rc = r1 ? r1 : r2;
    }
    if (!rc) {
      HIP_CHECK(hipEventRecord(s, st));
      for (int i = 0; i < reps && !rc; ++i) rc = hs_graph_replay(g, st);
      HIP_CHECK(hipEventRecord(e, st));
    }
    if (!rc) rc = hs_arena_sync(a, st);
    if (g) hs_graph_destroy(g);
    HIP_CHECK(hipStreamDestroy(st));
  }
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", hs_status_string(rc));
    return 1;
  }
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, s, e));
  printf("time (ms) is %f\n", ms / reps);
  int32_t sum = 0;
  HIP_CHECK(hipMemcpy(&sum, dout, sizeof(sum), hipMemcpyDeviceToHost));
  printf("Sum: %d\n", sum);
  printf("Correct sum: %d\n", correct);
  HIP_CHECK(hipFree(din));
  HIP_CHECK(hipFree(dpart));
  HIP_CHECK(hipFree(dout));
  hs_arena_destroy(a);
  return sum == correct ? 0 : 1;
}
