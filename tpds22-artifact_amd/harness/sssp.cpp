// sssp -- LONESTAR sssp on libhs: sssp [-q] [-o out] [-l] [-d delta] [-s startNode] graph.gr
// (LONESTAR/apps/sssp/support.cu: DELTA default 10000).
#include "lonestar.hpp"

int main(int argc, char **argv) {
  int c, quiet = 0, src = 0, delta = 10000;
  const char *out = nullptr;
  while ((c = getopt(argc, argv, "g:qo:ls:d:")) != -1) {
    switch (c) {
      case 'q': quiet = 1; break;
      case 'o': out = optarg; break;
      case 'g': case 'l': break;
      case 's': src = atoi(optarg); break;
      case 'd': delta = atoi(optarg); break;
      default:
        fprintf(stderr, "usage: %s [-q] [-o output-file] [-l] [-d delta] [-s startNode] graph-file\n", argv[0]);
        return EXIT_FAILURE;
    }
  }
  if (optind >= argc || delta <= 0) {
    fprintf(stderr, "usage: %s [-q] [-o output-file] [-l] [-d delta] [-s startNode] graph-file\n", argv[0]);
    return EXIT_FAILURE;
  }
  HostGraph g = read_graph(argv[optind], true);
  if (src < 0 || (uint32_t)src >= g.n) { fprintf(stderr, "ERROR: bad start node\n"); return EXIT_FAILURE; }
  int cus = 0;
  hs_arena *a = make_arena(&cus);
  const int bf = read_block_factor();
  printf("delta: %d\n", delta);
  DevGraph d;
  d.row = upload(g.row);
  d.col = upload(g.col);
  d.w = upload(g.w);
  std::vector<int32_t> dist(g.n);
  int32_t *dd = nullptr;
  void *work = nullptr;
  uint64_t wb = 0;
  uint32_t *stats = nullptr;
  hs_sssp_work_size(g.n, g.m, &wb);
  HIP_CHECK(hipMalloc(&dd, 4ull * g.n));
  HIP_CHECK(hipMalloc(&work, wb));
  HIP_CHECK(hipMalloc(&stats, 16));
  hipEvent_t s, e;
  HIP_CHECK(hipEventCreate(&s));
  HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipEventRecord(s, 0));
  int rc = hs_sssp(a, g.n, g.m, d.row, d.col, d.w, src, delta, dd, work, HS_BAR_XCD, bf > 0 ? bf * cus : 0,
                   stats, nullptr);
  HIP_CHECK(hipEventRecord(e, 0));
  if (!rc) rc = hs_arena_sync(a, nullptr);
  if (rc) die("hs_sssp", rc);
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, s, e));
  std::cout << "Kernel time (ms) " << ms << std::endl;
  uint32_t st[4];
  HIP_CHECK(hipMemcpy(st, stats, 16, hipMemcpyDeviceToHost));
  printf("iterations: %u\n", st[1]);
  HIP_CHECK(hipMemcpy(dist.data(), dd, 4ull * g.n, hipMemcpyDeviceToHost));
  if (!quiet) write_int_output(out, dist);
  hs_arena_destroy(a);
  return 0;
}
