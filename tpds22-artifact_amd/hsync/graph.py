"""Graph front end: CSR construction (R-MAT, .gr files) and the persistent
BFS / multi-source BFS / SSSP / PageRank kernels of libhs.

Mirrors the LonestarGPU apps' host flow (LONESTAR/skelapp/skel.cu:47-151:
read the graph, copy it to the device, run gg_main, write `"%d %d"` /
`"%d INF"` per node).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .api import Arena, _ptr, _stream, barrier_kind

INF = 2**31 - 1
SCHED = {"heuristic": 0, "push": 1, "pull": 2, "learned": 3, "forced": 4}
TRACE_WORDS = 16
MAX_LEVELS = 256


def _torch():
    import torch
    return torch


def rmat_csr(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19):
    lib = L.load()
    m = C.c_uint64(0)
    L.check(lib.hs_rmat_csr_host(scale, edge_factor, a, b, c, seed, None, None, C.byref(m)), "rmat")
    n = 1 << scale
    row = np.empty(n + 1, np.uint32)
    col = np.empty(m.value, np.uint32)
    L.check(lib.hs_rmat_csr_host(scale, edge_factor, a, b, c, seed, row.ctypes.data, col.ctypes.data,
                                 C.byref(m)), "rmat")
    return row, col


def weights(nedges: int, seed: int, lo: int = 1, hi: int = 255):
    w = np.empty(nedges, np.uint32)
    L.check(L.load().hs_gen_weights_host(w.ctypes.data, nedges, seed, lo, hi), "weights")
    return w


def transpose(row, col):
    n = row.size - 1
    rrow = np.empty_like(row)
    rcol = np.empty_like(col)
    L.check(L.load().hs_csr_transpose_host(n, col.size, row.ctypes.data, col.ctypes.data,
                                           rrow.ctypes.data, rcol.ctypes.data), "transpose")
    return rrow, rcol


def read_gr(path: str):
    """readFromGR (LONESTAR/rt/src/csr_graph.cu:152-237)."""
    lib = L.load()
    n, m, ety = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.check(lib.hs_gr_read_header(path.encode(), C.byref(n), C.byref(m), C.byref(ety)), path)
    row = np.empty(n.value + 1, np.uint32)
    col = np.empty(m.value, np.uint32)
    w = np.empty(m.value, np.uint32) if ety.value else None
    L.check(lib.hs_gr_read(path.encode(), row.ctypes.data, col.ctypes.data,
                           w.ctypes.data if w is not None else None), path)
    return row, col, w


def write_gr(path: str, row, col, w=None):
    L.check(L.load().hs_gr_write(path.encode(), row.size - 1, col.size, row.ctypes.data,
                                 col.ctypes.data, None if w is None else w.ctypes.data), path)


class Graph:
    """A CSR graph resident on the device (forward and transposed)."""

    def __init__(self, row, col, w=None, device="cuda", transposed=True):
        torch = _torch()
        self.n = int(row.size - 1)
        self.m = int(col.size)
        self.h_row, self.h_col, self.h_w = row, col, w
        t = lambda a: torch.from_numpy(a.view(np.int32)).to(device)  # noqa: E731
        self.row = t(row)
        self.col = t(col) if self.m else torch.zeros(1, dtype=torch.int32, device=device)
        self.w = t(w) if w is not None and self.m else None
        if transposed:
            rrow, rcol = transpose(row, col)
            self.h_rrow, self.h_rcol = rrow, rcol
            self.rrow = t(rrow)
            self.rcol = t(rcol) if self.m else torch.zeros(1, dtype=torch.int32, device=device)
        self.out_degree = np.diff(row.astype(np.int64))

    @staticmethod
    def rmat(scale, edge_factor=16, seed=1, weighted=False, wseed=None, **kw):
        row, col = rmat_csr(scale, edge_factor, seed)
        w = weights(col.size, seed + 1000 if wseed is None else wseed) if weighted else None
        return Graph(row, col, w, **kw)


def _work(nbytes: int, device="cuda"):
    torch = _torch()
    return torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=device)


def device_sources(srcs):
    """The sources as the launchers read them: a contiguous int32 device tensor.
    A tensor that already is one is used as it is (no host copy, which would
    wait for the stream's earlier work); a sequence is copied."""
    torch = _torch()
    if isinstance(srcs, torch.Tensor):
        if srcs.dtype != torch.int32 or not srcs.is_cuda or srcs.dim() != 1 or not srcs.is_contiguous():
            raise ValueError("sources tensor must be a contiguous 1-D int32 device tensor")
        return srcs
    return torch.tensor([int(x) for x in srcs], dtype=torch.int32, device="cuda")


def bfs(arena: Arena, g: Graph, src: int, barrier="xcd", num_wgs: int = 0, level=None,
        work=None, stream=None, sync: bool = True, lb: Optional[int] = None):
    """Single-source BFS; lb (the reference's -l): None = off, 0 = on with the
    reference's degree limit (the grid's thread count), else the degree limit."""
    torch = _torch()
    lib = L.load()
    nb = C.c_uint64(0)
    L.check(lib.hs_bfs_work_size(g.n, g.m, C.byref(nb)))
    work = work if work is not None else _work(nb.value)
    level = level if level is not None else torch.empty(g.n, dtype=torch.int32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = _stream(stream)
    if lb is None:
        L.check(lib.hs_bfs(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), src, _ptr(level), _ptr(work),
                           barrier_kind(barrier), num_wgs, _ptr(stats), s), "hs_bfs")
    else:
        L.check(lib.hs_bfs_lb(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), src, _ptr(level), _ptr(work),
                              barrier_kind(barrier), num_wgs, int(lb), _ptr(stats), s), "hs_bfs_lb")
    if sync:
        arena.sync(s)
    return level, stats


def msbfs(arena: Arena, g: Graph, srcs, schedule="heuristic", mlp=None,
          forced: Optional[Sequence[int]] = None, barrier="xcd", num_wgs: int = 0, levels=None,
          work=None, trace=None, stream=None, sync: bool = True):
    """Up to 64 sources per launch (a sequence, or an int32 device tensor:
    device_sources); returns levels [nsrc][n] (int32, INF unreached)."""
    torch = _torch()
    lib = L.load()
    nsrc = len(srcs)
    if not 0 < nsrc <= 64:
        raise ValueError("1..64 sources per launch")
    nb = C.c_uint64(0)
    L.check(lib.hs_msbfs_work_size(g.n, g.m, nsrc, C.byref(nb)))
    work = work if work is not None else _work(nb.value)
    levels = levels if levels is not None else torch.empty((nsrc, g.n), dtype=torch.int32, device="cuda")
    d_srcs = device_sources(srcs)
    sched = SCHED[schedule] if isinstance(schedule, str) else int(schedule)
    forced_buf = None
    if sched == 4:
        forced_buf = (C.c_uint8 * MAX_LEVELS)(*([int(x) for x in forced] + [0] * (MAX_LEVELS - len(forced))))
    s = _stream(stream)
    L.check(lib.hs_msbfs(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), _ptr(g.rrow), _ptr(g.rcol),
                         _ptr(d_srcs), nsrc, _ptr(levels), _ptr(work), barrier_kind(barrier), num_wgs,
                         sched, _ptr(mlp) if mlp is not None else None,
                         C.cast(forced_buf, C.c_void_p) if forced_buf is not None else None,
                         _ptr(trace) if trace is not None else None, s), "hs_msbfs")
    if sync:
        arena.sync(s)
    return levels


def new_trace():
    torch = _torch()
    return torch.zeros(MAX_LEVELS * TRACE_WORDS, dtype=torch.int32, device="cuda")


def decode_trace(trace) -> list:
    """per-level records: action (direction = bit 0), frontier size, frontier
    out-edges, open in-edges, discoveries, seconds between the level start and the
    end of its expansion"""
    t = trace.cpu().numpy().view(np.uint32).reshape(MAX_LEVELS, TRACE_WORDS).astype(np.uint64)
    out = []
    for r in t:
        start = int(r[7] | (r[8] << 32))
        end = int(r[9] | (r[10] << 32))
        if start == 0 and end == 0:
            break
        out.append({"act": int(r[0]), "dir": int(r[0]) & 1, "nf": int(r[1]), "mf": int(r[2] | (r[3] << 32)),
                    "mu": int(r[4] | (r[5] << 32)), "disc": int(r[6]),
                    "seconds": (end - start) / 100e6})  # s_memrealtime: 100 MHz
    return out


def sssp(arena: Arena, g: Graph, src: int, delta: int = 10000, barrier="xcd", num_wgs: int = 0,
         dist=None, work=None, stream=None, sync: bool = True):
    torch = _torch()
    lib = L.load()
    if g.w is None:
        raise ValueError("SSSP needs edge weights")
    nb = C.c_uint64(0)
    L.check(lib.hs_sssp_work_size(g.n, g.m, C.byref(nb)))
    work = work if work is not None else _work(nb.value)
    dist = dist if dist is not None else torch.empty(g.n, dtype=torch.int32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = _stream(stream)
    L.check(lib.hs_sssp(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), _ptr(g.w), src, delta,
                        _ptr(dist), _ptr(work), barrier_kind(barrier), num_wgs, _ptr(stats), s),
            "hs_sssp")
    if sync:
        arena.sync(s)
    return dist, stats


def mssp(arena: Arena, g: Graph, srcs, delta: int = 0x7FFFFFFF, barrier="xcd",
         num_wgs: int = 0, dist=None, work=None, stats=None, stream=None, sync: bool = True):
    """Up to 64 SSSP sources in one launch (a sequence, or an int32 device tensor:
    device_sources); returns dist [nsrc][n] (int32, INT32_MAX
    unreached).  delta: bucket width of the shared delta-stepping bound
    (INT32_MAX = frontier Bellman-Ford)."""
    torch = _torch()
    lib = L.load()
    if g.w is None:
        raise ValueError("SSSP needs edge weights")
    nsrc = len(srcs)
    if not 0 < nsrc <= 64:
        raise ValueError("1..64 sources per launch")
    nb = C.c_uint64(0)
    L.check(lib.hs_mssp_work_size(g.n, g.m, nsrc, C.byref(nb)))
    work = work if work is not None else _work(nb.value)
    dist = dist if dist is not None else torch.empty((nsrc, g.n), dtype=torch.int32, device="cuda")
    stats = stats if stats is not None else torch.zeros(4, dtype=torch.int32, device="cuda")
    d_srcs = device_sources(srcs)
    s = _stream(stream)
    L.check(lib.hs_mssp(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), _ptr(g.w), _ptr(d_srcs), nsrc,
                        delta, _ptr(dist), _ptr(work), barrier_kind(barrier), num_wgs, _ptr(stats), s),
            "hs_mssp")
    if sync:
        arena.sync(s)
    return dist, stats


def pagerank(arena: Arena, g: Graph, alpha=0.85, tol=1e-6, max_iter=1000, barrier="xcd",
             num_wgs: int = 0, rank=None, work=None, stream=None, sync: bool = True):
    torch = _torch()
    lib = L.load()
    nb = C.c_uint64(0)
    L.check(lib.hs_pagerank_work_size(g.n, g.m, C.byref(nb)))
    work = work if work is not None else _work(nb.value)
    rank = rank if rank is not None else torch.empty(g.n, dtype=torch.float32, device="cuda")
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = _stream(stream)
    L.check(lib.hs_pagerank_f32(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.rrow), _ptr(g.rcol),
                                alpha, tol, max_iter, _ptr(rank), _ptr(work), barrier_kind(barrier),
                                num_wgs, _ptr(stats), s), "hs_pagerank_f32")
    if sync:
        arena.sync(s)
    return rank, stats


def _tune_result():
    from .api import TuneResult
    return TuneResult()


def tune_msbfs(arena: Arena, g: Graph, srcs: Sequence[int], schedule="heuristic", mlp=None, reps: int = 3,
               levels=None, stream=None) -> dict:
    """hs_tune_msbfs: grid x barrier (fan-in) of the multi-source BFS on this graph"""
    torch = _torch()
    levels = levels if levels is not None else torch.empty((len(srcs), g.n), dtype=torch.int32, device="cuda")
    d_srcs = torch.tensor(list(srcs), dtype=torch.int32, device="cuda")
    sched = SCHED[schedule] if isinstance(schedule, str) else int(schedule)
    r = _tune_result()
    L.check(L.load().hs_tune_msbfs(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), _ptr(g.rrow), _ptr(g.rcol),
                                   _ptr(d_srcs), len(srcs), _ptr(levels), sched,
                                   _ptr(mlp) if mlp is not None else None, reps, C.byref(r), _stream(stream)),
            "hs_tune_msbfs")
    return r.as_dict()


def tune_bfs(arena: Arena, g: Graph, src: int, reps: int = 3, level=None, stream=None) -> dict:
    torch = _torch()
    level = level if level is not None else torch.empty(g.n, dtype=torch.int32, device="cuda")
    r = _tune_result()
    L.check(L.load().hs_tune_bfs(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), src, _ptr(level), reps,
                                 C.byref(r), _stream(stream)), "hs_tune_bfs")
    return r.as_dict()


def tune_mssp(arena: Arena, g: Graph, srcs: Sequence[int], delta: int = 0x7FFFFFFF, reps: int = 2, dist=None,
              stream=None) -> dict:
    torch = _torch()
    dist = dist if dist is not None else torch.empty((len(srcs), g.n), dtype=torch.int32, device="cuda")
    d_srcs = torch.tensor(list(srcs), dtype=torch.int32, device="cuda")
    r = _tune_result()
    L.check(L.load().hs_tune_mssp(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.col), _ptr(g.w), _ptr(d_srcs),
                                  len(srcs), delta, _ptr(dist), reps, C.byref(r), _stream(stream)), "hs_tune_mssp")
    return r.as_dict()


def tune_pagerank(arena: Arena, g: Graph, alpha=0.85, tol=1e-6, max_iter=1000, reps: int = 2, rank=None,
                  stream=None) -> dict:
    torch = _torch()
    rank = rank if rank is not None else torch.empty(g.n, dtype=torch.float32, device="cuda")
    r = _tune_result()
    L.check(L.load().hs_tune_pagerank_f32(arena.handle, g.n, g.m, _ptr(g.row), _ptr(g.rrow), _ptr(g.rcol),
                                          alpha, tol, max_iter, _ptr(rank), reps, C.byref(r), _stream(stream)),
            "hs_tune_pagerank_f32")
    return r.as_dict()


def reached_edges(levels, out_degree_t) -> int:
    """sum over sources of the out-degrees of the vertices each source reached
    (the traversed-edge count of the GTEPS metric)"""
    torch = _torch()
    reached = (levels != INF).to(torch.float64)
    return int((reached @ out_degree_t).sum().item())


def write_levels(path: str, level: np.ndarray) -> None:
    """output() of LONESTAR/apps/bfs/support.cu:27-46 ("%d %d" or "%d INF")."""
    with open(path, "w") as f:
        for i, v in enumerate(level.tolist()):
            f.write(f"{i} INF\n" if v == INF else f"{i} {v}\n")
