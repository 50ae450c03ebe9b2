// Host check of the serving loop's PageRank tile arithmetic
// (tpds22-artifact_amd/csrc/serve_tiles.hpp), built with g++ by
// tests/test_lib_cpu.py.  On random graphs (small in-degrees and hubs cut across
// many tiles) and random tile weights:
//  - phase 0's placement sets the start of every tile j in [1, nt) exactly once;
//  - the tiles' pieces cover every in-edge exactly once, and every vertex is
//    either whole in exactly one tile or cut (then only in pieces);
//  - with arbitrary tile starts (what a tile could read before phase 0's writes
//    are visible) every vertex index stays below n and every in-edge range
//    inside its vertex's range.
// "plant" as the first argument shifts the start of every head piece by one
// in-edge in this program's copy of the sums, and the check must then fail.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "serve_tiles.hpp"

using hs_serve::PrPiece;
using hs_serve::PrTileSpan;

int main(int argc, char **argv) {
  const bool plant = argc > 1 && !strcmp(argv[1], "plant");
  std::mt19937_64 rng(20261203);
  long fails = 0, tiles_checked = 0, cut_vertices = 0;
  for (int trial = 0; trial < 400; ++trial) {
    const uint32_t n = 1 + (uint32_t)(rng() % 3000);
    std::vector<uint32_t> rrow(n + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
      const uint32_t d = rng() % 16 == 0 ? (uint32_t)(rng() % 6000) : (uint32_t)(rng() % 9);
      rrow[v + 1] = rrow[v] + d;
    }
    const uint64_t W = rrow[n] + 4ull * n;
    const uint64_t wt = 1 + rng() % (trial % 2 ? 300 : 20000);
    const uint32_t nt = (uint32_t)((W + wt - 1) / wt);
    std::vector<uint32_t> bv(nt + 1, 0), bo(nt + 1, 0), placed(nt + 1, 0);
    for (uint32_t v = 0; v < n; ++v)
      hs_serve::pr_place(v, rrow[v], rrow[v + 1], wt, nt, [&](uint32_t j, uint32_t off) {
        bv[j] = v;
        bo[j] = off;
        ++placed[j];
      });
    for (uint32_t j = 1; j < nt; ++j)
      if (placed[j] != 1) ++fails;
    std::vector<uint32_t> covered(rrow[n], 0), whole(n, 0), pieces(n, 0);
    for (uint32_t t = 0; t < nt; ++t) {
      const bool last = t + 1 >= nt;
      const PrTileSpan sp = hs_serve::pr_tile_span(t, nt, n, bv[t], bo[t], last ? 0u : bv[t + 1], last ? 0u : bo[t + 1]);
      ++tiles_checked;
      if (sp.vend > n) ++fails;
      for (uint32_t v = sp.v0; v < sp.vend && v < n; ++v) {
        PrPiece pc = hs_serve::pr_piece(sp, v, rrow[v], rrow[v + 1]);
        if (plant && pc.cut_head && pc.es < pc.ee) ++pc.es;
        if (pc.es < rrow[v] || pc.es > pc.ee || pc.ee > rrow[v + 1]) { ++fails; continue; }
        for (uint32_t e = pc.es; e < pc.ee; ++e) ++covered[e];
        if (pc.cut_head || pc.cut_tail) ++pieces[v];
        else ++whole[v];
      }
    }
    for (uint32_t e = 0; e < rrow[n]; ++e)
      if (covered[e] != 1) ++fails;
    for (uint32_t v = 0; v < n; ++v) {
      if (pieces[v]) ++cut_vertices;
      if (pieces[v] ? whole[v] != 0 || pieces[v] < 2 : whole[v] != 1) ++fails;
    }
    // arbitrary starts: indices stay in range
    for (uint32_t t = 0; t < 64; ++t) {
      const uint32_t ntx = 1 + (uint32_t)(rng() % 100000);
      const uint32_t tx = (uint32_t)(rng() % ntx);
      const PrTileSpan sp = hs_serve::pr_tile_span(tx, ntx, n, (uint32_t)rng(), (uint32_t)rng(), (uint32_t)rng(),
                                                   (uint32_t)rng());
      if (sp.v0 >= n || sp.vend > n) ++fails;
      const uint32_t vb = sp.v0, ve = sp.vend < vb + 256 ? sp.vend : vb + 256;
      for (uint32_t v = vb; v < ve; ++v) {
        const PrPiece pc = hs_serve::pr_piece(sp, v, rrow[v], rrow[v + 1]);
        if (pc.es < rrow[v] || pc.es > pc.ee || pc.ee > rrow[v + 1]) ++fails;
      }
      if (sp.vend > vb) {  // the tile's last vertex too
        const uint32_t v = sp.vend - 1;
        const PrPiece pc = hs_serve::pr_piece(sp, v, rrow[v], rrow[v + 1]);
        if (pc.es < rrow[v] || pc.es > pc.ee || pc.ee > rrow[v + 1]) ++fails;
      }
    }
  }
  printf("tiles %ld cut vertices %ld failures %ld\n", tiles_checked, cut_vertices, fails);
  return fails ? 1 : 0;
}
