// capture.hip -- hipGraph capture and replay of repeated launches.
//
// The reference repeats a launch in a host loop (Heterosync/main.cu:1507-1526,
// NUM_REPEATS) and pays the launch path every time.  Here a sequence of libhs
// launches on one stream is captured once into a hipGraph and replayed as one
// submission.  Every libhs launcher is capturable: device words are cleared by
// fill kernels, never by hipMemsetAsync, and nothing allocates between the
// launches of a captured sequence.
#include <hip/hip_runtime.h>

#include <stdlib.h>

#include "hs_internal.hpp"

struct hs_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

int hs_graph_capture_begin(void *stream) {
  // the legacy null stream cannot be captured
  if (!stream) return HS_ERR_INVALID;
  HS_TRY_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return HS_OK;
}

int hs_graph_capture_end(void *stream, hs_graph **out) {
  if (!stream || !out) return HS_ERR_INVALID;
  *out = nullptr;
  hipGraph_t g = nullptr;
  HS_TRY_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  if (!g) return HS_ERR_HIP;
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  if (rc != hipSuccess) {
    fprintf(stderr, "libhs: hipGraphInstantiate failed: %s\n", hipGetErrorString(rc));
    (void)hipGraphDestroy(g);  // already failing
    return HS_ERR_HIP;
  }
  hs_graph *h = (hs_graph *)calloc(1, sizeof(hs_graph));
  if (!h) {
    (void)hipGraphExecDestroy(e);
    (void)hipGraphDestroy(g);
    return HS_ERR_NOMEM;
  }
  h->graph = g;
  h->exec = e;
  *out = h;
  return HS_OK;
}

int hs_graph_replay(hs_graph *g, void *stream) {
  if (!g) return HS_ERR_INVALID;
  HS_TRY_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return HS_OK;
}

int hs_graph_destroy(hs_graph *g) {
  if (!g) return HS_ERR_INVALID;
  const hipError_t a = hipGraphExecDestroy(g->exec);
  const hipError_t b = hipGraphDestroy(g->graph);
  free(g);
  return (a == hipSuccess && b == hipSuccess) ? HS_OK : HS_ERR_HIP;
}

}  // extern "C"
