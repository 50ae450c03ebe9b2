// gen.hip -- seeded synthetic inputs and the graph formats.
//
// Inputs come from a counter-based SplitMix64 stream (Steele, Lea & Flood 2014;
// the constants Vigna publishes): element i of stream `seed` is
// mix64(mix64(seed) + (i+1) * 0x9E3779B97F4A7C15).  Being counter based, the
// same values are produced by the device kernel, by the host function and by any
// restatement of it, for any thread count.  (The reference feeds all-ones
// arrays, Reduce/test_reduce.cu:14 and scan/main.cu:39, and downloads its
// graphs, LONESTAR/Makefile:25-31; it has no generator.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdio.h>
#include <string.h>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "hs_internal.hpp"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t rand64(uint64_t seed_state, uint64_t i) {
  return mix64(seed_state + (i + 1) * kGolden);
}

// Lemire multiply-shift map of the high 32 bits onto [lo, hi]
__host__ __device__ __forceinline__ uint32_t to_range(uint64_t r, uint32_t lo, uint64_t range) {
  return lo + (uint32_t)(((r >> 32) * range) >> 32);
}

__global__ void k_gen_uniform(uint32_t *out, uint64_t n, uint64_t state, uint32_t lo, uint64_t range) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = to_range(rand64(state, i), lo, range);
}

}  // namespace

extern "C" {

int hs_gen_uniform_u32(uint32_t *d_out, uint64_t n, uint64_t seed, uint32_t lo, uint32_t hi, void *stream) {
  if ((n && !d_out) || hi < lo) return HS_ERR_INVALID;
  if (!n) return HS_OK;
  const uint64_t range = (uint64_t)hi - lo + 1;
  const int blocks = (int)std::min<uint64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_gen_uniform, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d_out, n,
                     mix64(seed), lo, range);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

int hs_gen_uniform_u32_host(uint32_t *h_out, uint64_t n, uint64_t seed, uint32_t lo, uint32_t hi) {
  if ((n && !h_out) || hi < lo) return HS_ERR_INVALID;
  const uint64_t range = (uint64_t)hi - lo + 1, st = mix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) h_out[i] = to_range(rand64(st, (uint64_t)i), lo, range);
  return HS_OK;
}

// R-MAT (Chakrabarti, Zhan & Faloutsos 2004).  Edge e draws, for each of the
// `scale` levels from the most significant bit down, one uniform double
// u = (rand64(state, e*scale + level) >> 11) * 2^-53 and picks the quadrant
// (0,0) if u < a, (0,1) if u < a+b, (1,0) if u < a+b+c, else (1,1).  The CSR keeps
// edges as generated (directed, vertex ids unpermuted), drops self loops and
// merges duplicates; neighbour lists are ascending.
int hs_rmat_csr_host(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                     uint32_t *h_row, uint32_t *h_col, uint64_t *nedges_out) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || !nedges_out || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return HS_ERR_INVALID;
  const uint64_t n = 1ull << scale, m = n * (uint64_t)edge_factor;
  const uint64_t st = mix64(seed);
  const double ab = a + b, abc = a + b + c;
  std::vector<uint64_t> edges(m);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m; ++e) {
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      const double r = (double)(rand64(st, (uint64_t)e * scale + l) >> 11) * (1.0 / 9007199254740992.0);
      const uint64_t ub = r >= ab, vb = (r >= a && r < ab) || r >= abc;
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    edges[e] = (u << 32) | v;
  }
  // counting sort by source (stable; per-thread histograms)
  std::vector<uint64_t> deg(n + 1, 0);
  for (uint64_t e = 0; e < m; ++e) {
    const uint64_t u = edges[e] >> 32, v = edges[e] & 0xFFFFFFFFu;
    if (u != v) deg[u + 1]++;
  }
  for (uint64_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<uint32_t> col(deg[n]);
  {
    std::vector<uint64_t> pos(deg.begin(), deg.end() - 1);
    for (uint64_t e = 0; e < m; ++e) {
      const uint64_t u = edges[e] >> 32, v = edges[e] & 0xFFFFFFFFu;
      if (u != v) col[pos[u]++] = (uint32_t)v;
    }
  }
  std::vector<uint64_t>().swap(edges);
  // per-row sort + unique
  std::vector<uint64_t> ucount(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    uint32_t *bgn = col.data() + deg[u], *end = col.data() + deg[u + 1];
    std::sort(bgn, end);
    ucount[u] = (uint64_t)(std::unique(bgn, end) - bgn);
  }
  uint64_t total = 0;
  for (uint64_t u = 0; u < n; ++u) total += ucount[u];
  *nedges_out = total;
  if (total > 0xFFFFFFFFull) return HS_ERR_INVALID;  // uint32 CSR offsets
  if (!h_row || !h_col) return HS_OK;
  h_row[0] = 0;
  for (uint64_t u = 0; u < n; ++u) h_row[u + 1] = (uint32_t)(h_row[u] + ucount[u]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; ++u)
    memcpy(h_col + h_row[u], col.data() + deg[u], sizeof(uint32_t) * ucount[u]);
  return HS_OK;
}

// SSSP weights: edge i of the CSR gets uniform [lo, hi] from stream `seed`
int hs_gen_weights_host(uint32_t *h_w, uint64_t nedges, uint64_t seed, uint32_t lo, uint32_t hi) {
  return hs_gen_uniform_u32_host(h_w, nedges, seed, lo, hi);
}

int hs_csr_transpose_host(uint32_t nnodes, uint64_t nedges, const uint32_t *row, const uint32_t *col,
                          uint32_t *rrow, uint32_t *rcol) {
  if (!row || !col || !rrow || !rcol) return HS_ERR_INVALID;
  std::vector<uint32_t> cnt((size_t)nnodes + 1, 0);
  for (uint64_t e = 0; e < nedges; ++e) cnt[col[e] + 1]++;
  rrow[0] = 0;
  for (uint32_t v = 0; v < nnodes; ++v) rrow[v + 1] = rrow[v] + cnt[v + 1];
  std::vector<uint32_t> pos(rrow, rrow + nnodes);
  // sources visited in ascending order -> each reverse list is ascending
  for (uint32_t u = 0; u < nnodes; ++u)
    for (uint32_t e = row[u]; e < row[u + 1]; ++e) rcol[pos[col[e]]++] = u;
  return HS_OK;
}

// ---- .gr v1 (LONESTAR/rt/src/csr_graph.cu:152-237) ----
// u64 version(=1), u64 edge-type size, u64 nnodes, u64 nedges, u64 row_end[nnodes],
// u32 dst[nedges], 4 bytes of padding when nedges is odd, u32 data[nedges] when
// the edge-type size is non-zero.  Little-endian.
int hs_gr_read_header(const char *path, uint64_t *nnodes, uint64_t *nedges, uint64_t *edge_ty_size) {
  if (!path || !nnodes || !nedges || !edge_ty_size) return HS_ERR_INVALID;
  FILE *f = fopen(path, "rb");
  if (!f) return HS_ERR_INVALID;
  uint64_t h[4];
  const size_t got = fread(h, 8, 4, f);
  fclose(f);
  if (got != 4 || h[0] != 1) return HS_ERR_INVALID;
  *edge_ty_size = h[1];
  *nnodes = h[2];
  *nedges = h[3];
  return HS_OK;
}

int hs_gr_read(const char *path, uint32_t *row, uint32_t *col, uint32_t *w) {
  if (!path || !row || !col) return HS_ERR_INVALID;
  FILE *f = fopen(path, "rb");
  if (!f) return HS_ERR_INVALID;
  uint64_t h[4];
  int rc = HS_OK;
  if (fread(h, 8, 4, f) != 4 || h[0] != 1 || h[2] > 0xFFFFFFFFull || h[3] > 0xFFFFFFFFull) rc = HS_ERR_INVALID;
  const uint64_t n = h[2], m = h[3];
  if (!rc) {
    std::vector<uint64_t> ends(n);
    if (fread(ends.data(), 8, n, f) != n) rc = HS_ERR_INVALID;
    row[0] = 0;
    for (uint64_t v = 0; !rc && v < n; ++v) {
      if (ends[v] > m || ends[v] < row[v]) rc = HS_ERR_INVALID;
      else row[v + 1] = (uint32_t)ends[v];
    }
  }
  if (!rc && fread(col, 4, m, f) != m) rc = HS_ERR_INVALID;
  for (uint64_t e = 0; !rc && e < m; ++e)
    if (col[e] >= n) rc = HS_ERR_INVALID;  // csr_graph.cu:213 reports invalid edges
  if (!rc && h[1] && w) {
    uint32_t pad;
    if (m % 2 && fread(&pad, 4, 1, f) != 1) rc = HS_ERR_INVALID;
    if (!rc && h[1] != 4) rc = HS_ERR_INVALID;
    if (!rc && fread(w, 4, m, f) != m) rc = HS_ERR_INVALID;
  }
  fclose(f);
  return rc;
}

int hs_gr_write(const char *path, uint32_t nnodes, uint64_t nedges, const uint32_t *row, const uint32_t *col,
                const uint32_t *w) {
  if (!path || !row || (nedges && !col)) return HS_ERR_INVALID;
  FILE *f = fopen(path, "wb");
  if (!f) return HS_ERR_INVALID;
  const uint64_t h[4] = {1, w ? 4u : 0u, nnodes, nedges};
  bool ok = fwrite(h, 8, 4, f) == 4;
  for (uint32_t v = 0; ok && v < nnodes; ++v) {
    const uint64_t e = row[v + 1];
    ok = fwrite(&e, 8, 1, f) == 1;
  }
  if (ok && nedges) ok = fwrite(col, 4, nedges, f) == nedges;
  if (ok && w) {
    const uint32_t z = 0;
    if (nedges % 2) ok = fwrite(&z, 4, 1, f) == 1;
    if (ok && nedges) ok = fwrite(w, 4, nedges, f) == nedges;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? HS_OK : HS_ERR_INVALID;
}

}  // extern "C"
