// lonestar.hpp -- shared command-line harness of the graph apps (bfs, sssp, pr),
// the shape of LONESTAR/skelapp/skel.cu: getopt "-q -g gpu -o output-file" plus
// the app's options, one positional .gr file, the persistent grid's block factor
// read from stdin ("Enter Block factor", kernel_gsrb.cu:716-719; a non-positive
// or missing value keeps the default grid), timing of gg_main, output file.
#pragma once

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "hip_check.hpp"
#include "hs.h"
#include "hs_graph.h"

struct HostGraph {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> row, col, w;
};

struct DevGraph {
  uint32_t *row = nullptr, *col = nullptr, *w = nullptr, *rrow = nullptr, *rcol = nullptr;
};

inline void die(const char *what, int rc) {
  fprintf(stderr, "ERROR: %s: %s\n", what, hs_status_string(rc));
  exit(EXIT_FAILURE);
}

inline HostGraph read_graph(const char *path, bool want_weights) {
  HostGraph g;
  uint64_t n = 0, m = 0, ety = 0;
  int rc = hs_gr_read_header(path, &n, &m, &ety);
  if (rc) die(path, rc);
  printf("nnodes=%llu, nedges=%llu.\n", (unsigned long long)n, (unsigned long long)m);
  g.n = (uint32_t)n;
  g.m = m;
  g.row.resize(n + 1);
  g.col.resize(m ? m : 1);
  if (ety) g.w.resize(m ? m : 1);
  rc = hs_gr_read(path, g.row.data(), g.col.data(), ety ? g.w.data() : nullptr);
  if (rc) die(path, rc);
  if (want_weights && !ety) {
    fprintf(stderr, "ERROR: %s has no edge data\n", path);
    exit(EXIT_FAILURE);
  }
  return g;
}

template <class T>
inline T *upload(const std::vector<T> &v) {
  T *d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (v.empty() ? 1 : v.size())) != hipSuccess) die("hipMalloc", HS_ERR_NOMEM);
  if (!v.empty()) HIP_CHECK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return d;
}

inline int read_block_factor() {
  std::cout << " Enter Block factor" << std::endl;
  int bf = 0;
  if (!(std::cin >> bf)) bf = 0;
  return bf;
}

inline hs_arena *make_arena(int *num_cus) {
  hs_device_info info;
  int rc = hs_device_info_get(&info);
  if (rc) die("hs_device_info_get", rc);
  printf("Using GPU: %s (%d CUs, %d XCDs)\n", info.arch, info.num_cus, info.num_xcds);
  *num_cus = info.num_cus;
  hs_arena_desc d = {8 * info.num_cus, 1, 1, info.num_cus, 0, 0};
  hs_arena *a = nullptr;
  rc = hs_arena_create(&d, &a);
  if (rc) die("hs_arena_create", rc);
  return a;
}

inline void write_int_output(const char *path, const std::vector<int32_t> &v) {
  if (!path) return;
  FILE *f = strcmp(path, "-") == 0 ? stdout : fopen(path, "w");
  if (!f) {
    fprintf(stderr, "ERROR: cannot open %s\n", path);
    exit(EXIT_FAILURE);
  }
  for (size_t i = 0; i < v.size(); ++i) {
    if (v[i] == 0x7FFFFFFF) fprintf(f, "%zu INF\n", i);
    else fprintf(f, "%zu %d\n", i, v[i]);
  }
  if (f != stdout) fclose(f);
}
