// pr -- LONESTAR pr on libhs: pr [-q] [-o out] [-n] [-t top] [-x max_iterations] graph.gr
// (LONESTAR/apps/pr/support.cu).  Ranks are the normalized PageRank (sum 1,
// damping 0.85, L1 change < 1e-6), printed "%d %f" per node; -n prints node ids
// only; -t prints the top ranks.  "PR took N iterations" as kernel_gsrb.cu:712.
#include <algorithm>
#include <numeric>

#include "lonestar.hpp"

int main(int argc, char **argv) {
  int c, quiet = 0, no_rank = 0, top = 0, max_iter = 1000000;
  const char *out = nullptr;
  while ((c = getopt(argc, argv, "g:qo:nt:x:")) != -1) {
    switch (c) {
      case 'q': quiet = 1; break;
      case 'o': out = optarg; break;
      case 'g': break;
      case 'n': no_rank = 1; break;
      case 't': top = atoi(optarg); break;
      case 'x': max_iter = atoi(optarg); break;
      default:
        fprintf(stderr, "usage: %s [-q] [-o output-file] [-n] [-t top_ranks] [-x max_iterations] graph-file\n", argv[0]);
        return EXIT_FAILURE;
    }
  }
  if (optind >= argc || max_iter < 1) {
    fprintf(stderr, "usage: %s [-q] [-o output-file] [-n] [-t top_ranks] [-x max_iterations] graph-file\n", argv[0]);
    return EXIT_FAILURE;
  }
  HostGraph g = read_graph(argv[optind], false);
  int cus = 0;
  hs_arena *a = make_arena(&cus);
  const int bf = read_block_factor();
  std::vector<uint32_t> rrow(g.n + 1), rcol(g.m ? g.m : 1);
  hs_csr_transpose_host(g.n, g.m, g.row.data(), g.col.data(), rrow.data(), rcol.data());
  DevGraph d;
  d.row = upload(g.row);
  d.rrow = upload(rrow);
  d.rcol = upload(rcol);
  float *dr = nullptr;
  void *work = nullptr;
  uint64_t wb = 0;
  uint32_t *stats = nullptr;
  hs_pagerank_work_size(g.n, g.m, &wb);
  HIP_CHECK(hipMalloc(&dr, 4ull * g.n));
  HIP_CHECK(hipMalloc(&work, wb));
  HIP_CHECK(hipMalloc(&stats, 16));
  hipEvent_t s, e;
  HIP_CHECK(hipEventCreate(&s));
  HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipEventRecord(s, 0));
  int rc = hs_pagerank_f32(a, g.n, g.m, d.row, d.rrow, d.rcol, 0.85f, 1e-6f, max_iter, dr, work, HS_BAR_XCD,
                           bf > 0 ? bf * cus : 0, stats, nullptr);
  HIP_CHECK(hipEventRecord(e, 0));
  if (!rc) rc = hs_arena_sync(a, nullptr);
  if (rc) die("hs_pagerank_f32", rc);
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, s, e));
  std::cout << "Kernel time (ms) " << ms << std::endl;
  uint32_t st[4];
  HIP_CHECK(hipMemcpy(st, stats, 16, hipMemcpyDeviceToHost));
  printf("PR took %u iterations\n", st[0]);
  std::vector<float> rank(g.n);
  HIP_CHECK(hipMemcpy(rank.data(), dr, 4ull * g.n, hipMemcpyDeviceToHost));
  double sum = 0;
  for (float r : rank) sum += r;
  printf("sum: %f (%u)\n", sum, g.n);
  if (!quiet && out) {
    FILE *f = strcmp(out, "-") == 0 ? stdout : fopen(out, "w");
    if (!f) { fprintf(stderr, "ERROR: cannot open %s\n", out); return EXIT_FAILURE; }
    std::vector<uint32_t> order(g.n);
    std::iota(order.begin(), order.end(), 0u);
    if (top > 0) {
      std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return rank[x] > rank[y]; });
      order.resize(std::min<size_t>(top, order.size()));
    }
    for (size_t i = 0; i < order.size(); ++i) {
      if (no_rank) fprintf(f, "%u\n", order[i]);
      else fprintf(f, "%u %.9g\n", order[i], rank[order[i]]);
    }
    if (f != stdout) fclose(f);
  }
  hs_arena_destroy(a);
  return 0;
}
