// bfs -- LONESTAR bfs on libhs: bfs [-q] [-o out] [-l] [-s startNode] graph.gr
// (LONESTAR/apps/bfs/support.cu, LONESTAR/skelapp/skel.cu).  -l turns on the
// load balancer (hs_bfs_lb: frontier vertices of out-degree >= the grid's
// thread count have their edges split evenly over the grid, as the reference's
// enable_lb path, LONESTAR/apps/bfs/kernel_gsrb.cu:569-579).
#include "lonestar.hpp"

int main(int argc, char **argv) {
  int c, quiet = 0, src = 0, lb = 0;
  const char *out = nullptr;
  while ((c = getopt(argc, argv, "g:qo:ls:")) != -1) {
    switch (c) {
      case 'q': quiet = 1; break;
      case 'o': out = optarg; break;
      case 'g': break;
      case 'l': lb = 1; break;
      case 's': src = atoi(optarg); break;
      default:
        fprintf(stderr, "usage: %s [-q] [-g gpunum] [-o output-file] [-l] [-s startNode] graph-file\n", argv[0]);
        return EXIT_FAILURE;
    }
  }
  if (optind >= argc) {
    fprintf(stderr, "usage: %s [-q] [-o output-file] [-l] [-s startNode] graph-file\n", argv[0]);
    return EXIT_FAILURE;
  }
  HostGraph g = read_graph(argv[optind], false);
  if (src < 0 || (uint32_t)src >= g.n) { fprintf(stderr, "ERROR: bad start node\n"); return EXIT_FAILURE; }
  int cus = 0;
  hs_arena *a = make_arena(&cus);
  const int bf = read_block_factor();
  DevGraph d;
  d.row = upload(g.row);
  d.col = upload(g.col);
  std::vector<int32_t> level(g.n);
  int32_t *dl = nullptr;
  void *work = nullptr;
  uint64_t wb = 0;
  uint32_t *stats = nullptr;
  hs_bfs_work_size(g.n, g.m, &wb);
  HIP_CHECK(hipMalloc(&dl, 4ull * g.n));
  HIP_CHECK(hipMalloc(&work, wb));
  HIP_CHECK(hipMalloc(&stats, 16));
  hipEvent_t s, e;
  HIP_CHECK(hipEventCreate(&s));
  HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipEventRecord(s, 0));
  const int wgs = bf > 0 ? bf * cus : 0;
  int rc = lb ? hs_bfs_lb(a, g.n, g.m, d.row, d.col, src, dl, work, HS_BAR_XCD, wgs, 0, stats, nullptr)
              : hs_bfs(a, g.n, g.m, d.row, d.col, src, dl, work, HS_BAR_XCD, wgs, stats, nullptr);
  HIP_CHECK(hipEventRecord(e, 0));
  if (!rc) rc = hs_arena_sync(a, nullptr);
  if (rc) die("hs_bfs", rc);
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, s, e));
  std::cout << "Kernel time (ms) " << ms << std::endl;
  uint32_t st[4];
  HIP_CHECK(hipMemcpy(st, stats, 16, hipMemcpyDeviceToHost));
  printf("levels: %u\n", st[0]);
  HIP_CHECK(hipMemcpy(level.data(), dl, 4ull * g.n, hipMemcpyDeviceToHost));
  if (!quiet) write_int_output(out, level);
  hs_arena_destroy(a);
  return 0;
}
