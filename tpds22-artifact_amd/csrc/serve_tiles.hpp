// serve_tiles.hpp -- the index arithmetic of the serving loop's PageRank
// iteration tiles, shared by the kernel (serve.hip) and a host check
// (tests/test_lib_cpu.py builds it with g++ and runs it on random graphs and
// on arbitrary tile starts).
//
// Iteration tiles are ranges of equal weight in the sequence of vertices, each
// vertex weighing 4 plus its in-degree: position F(v) = rrow[v] + 4 v starts
// vertex v, tile j starts at position j wt.  Phase 0 records, for every tile
// j >= 1, the vertex bv[j] whose span [F(v), F(v+1)) holds j wt and the in-edge
// offset bo[j] of the boundary inside that vertex (0 when the boundary falls
// on the vertex's 4 weight units, i.e. before its first in-edge).  A boundary
// with bo[j] > 0 cuts vertex bv[j]: tile j-1 sums its in-edges [0, bo[j]) (the
// tail piece) and tile j the rest, or up to the next boundary (head piece).
// Every index a tile derives from (bv, bo) is clamped into the graph, so tile
// starts read before they are visible give wrong sums, never an access out of
// range.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define HS_TILE_FN __host__ __device__ inline
#else
#define HS_TILE_FN inline
#endif

namespace hs_serve {

// the tiles j in [1, nt) whose start j * wt vertex v holds, with the in-edge
// offset of each: calls put(j, bo) for them in increasing j
template <class Put>
HS_TILE_FN void pr_place(uint32_t v, uint32_t rb, uint32_t re, uint64_t wt, uint32_t nt, Put put) {
  const uint64_t f0 = (uint64_t)rb + 4ull * v, f1 = (uint64_t)re + 4ull * (v + 1);
  for (uint64_t j = (f0 + wt - 1) / wt; j * wt < f1 && j < nt; ++j) {
    if (!j) continue;
    const uint64_t pos = j * wt;
    put((uint32_t)j, pos >= f0 + 4 ? (uint32_t)(pos - f0 - 4) : 0u);
  }
}

// the vertices of tile t: [v0, vend), where v0 may be cut at in-edge offset o0
// (o0 > 0: the tile holds v0's in-edges from o0 on) and v1 = vend - 1 may be
// cut at o1 (the tile holds v1's in-edges before o1) unless the tile is last
struct PrTileSpan {
  uint32_t v0, o0, v1, o1, vend;
  bool last;
};

HS_TILE_FN PrTileSpan pr_tile_span(uint32_t t, uint32_t nt, uint32_t n, uint32_t bv_t, uint32_t bo_t,
                                   uint32_t bv_next, uint32_t bo_next) {
  PrTileSpan s;
  s.last = t + 1 >= nt;
  s.v0 = t ? (bv_t < n - 1 ? bv_t : n - 1) : 0u;
  s.o0 = t ? bo_t : 0u;
  s.v1 = s.last ? n : (bv_next < n - 1 ? bv_next : n - 1);
  s.o1 = s.last ? 0u : bo_next;
  s.vend = (!s.last && s.o1) ? s.v1 + 1 : s.v1;  // v1 is in the tile only when cut
  return s;
}

// the in-edges [es, ee) of vertex v (in-edge range [rb, re)) that the tile sums,
// and whether they are v's head or tail piece
struct PrPiece {
  uint32_t es, ee;
  bool cut_head, cut_tail;
};

// rb + off clamped to re, without wrapping past 2^32 (off may be arbitrary)
HS_TILE_FN uint32_t pr_edge_at(uint32_t rb, uint32_t re, uint32_t off) { return off < re - rb ? rb + off : re; }

HS_TILE_FN PrPiece pr_piece(const PrTileSpan &s, uint32_t v, uint32_t rb, uint32_t re) {
  PrPiece p;
  p.cut_head = v == s.v0 && s.o0 > 0;
  p.cut_tail = !s.last && v == s.v1 && s.o1 > 0;
  p.es = pr_edge_at(rb, re, v == s.v0 ? s.o0 : 0u);
  if (p.cut_tail) {
    const uint32_t b = pr_edge_at(rb, re, s.o1);
    p.ee = b > p.es ? b : p.es;
  } else {
    p.ee = re;
  }
  return p;
}

}  // namespace hs_serve
