"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- the checker.

Test infrastructure only; the product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_P = C.c_void_p
_lib = None


def load():
    global _lib
    if _lib is not None:
        return Oracle(_lib)
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = C.CDLL(ORACLE_SO)
    sig = {
        "orc_mix64": (C.c_uint64, [C.c_uint64]),
        "orc_rand64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_gen_uniform_u32": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _P]),
        "orc_reduce_i32": (C.c_int32, [_P, C.c_uint64]),
        "orc_reduce_i64": (C.c_int64, [_P, C.c_uint64]),
        "orc_scan_exclusive_u32": (None, [_P, C.c_uint64, _P]),
        "orc_scan_inclusive_u32": (None, [_P, C.c_uint64, _P]),
        "orc_hs_prim_id": (C.c_int, [C.c_char_p]),
        "orc_hs_num_locs_mult": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "orc_hs_golden": (None, [C.c_int] * 7 + [_P, C.c_int64]),
        "orc_bfs": (None, [C.c_int32, _P, _P, C.c_int32, _P]),
        "orc_sssp": (None, [C.c_int32, _P, _P, _P, C.c_int32, _P]),
        "orc_pagerank_f64": (C.c_int, [C.c_int32, _P, _P, C.c_double, C.c_double, C.c_int, _P]),
        "orc_pagerank_push_fixpoint_f64": (C.c_int, [C.c_int32, _P, _P, C.c_double, C.c_double,
                                                     C.c_int, _P]),
        "orc_gr_write": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint64, _P, _P, _P]),
        "orc_gr_read_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64)]),
        "orc_gr_read": (C.c_int, [C.c_char_p, _P, _P, _P]),
        "orc_par_reduce_i64": (C.c_int64, [_P, C.c_uint64, C.POINTER(C.c_int)]),
        "orc_par_scan_inclusive_u32": (None, [_P, C.c_uint64, _P, C.POINTER(C.c_int)]),
        "orc_cpu_barrier_ns": (C.c_double, [C.c_int, C.POINTER(C.c_int)]),
        "orc_cpu_msbfs": (None, [C.c_int32, _P, _P, _P, _P, _P, C.c_int, _P, C.POINTER(C.c_int)]),
        "orc_cpu_mssp": (None, [C.c_int32, _P, _P, _P, _P, C.c_int, _P, C.POINTER(C.c_int)]),
        "orc_cpu_pagerank_pull_f32": (C.c_int, [C.c_int32, _P, _P, _P, C.c_double, C.c_double, C.c_int,
                                                _P, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return Oracle(lib)


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """numpy-friendly wrappers around the C oracle."""

    def __init__(self, lib):
        self.lib = lib

    def mix64(self, x):
        return self.lib.orc_mix64(x)

    def rand64(self, seed, i):
        return self.lib.orc_rand64(seed, i)

    def gen_uniform_u32(self, seed, n, lo, hi):
        out = np.empty(n, dtype=np.uint32)
        self.lib.orc_gen_uniform_u32(seed, n, lo, hi, _p(out))
        return out

    def reduce_i32(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        return int(self.lib.orc_reduce_i32(_p(a), a.size))

    def reduce_i64(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        return int(self.lib.orc_reduce_i64(_p(a), a.size))

    def scan_u32(self, a, inclusive):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.empty_like(a)
        f = self.lib.orc_scan_inclusive_u32 if inclusive else self.lib.orc_scan_exclusive_u32
        f(_p(a), a.size, _p(out))
        return out

    def prim_id(self, name):
        return self.lib.orc_hs_prim_id(name.encode())

    def num_locs_mult(self, prim, num_tbs, num_sm):
        return self.lib.orc_hs_num_locs_mult(prim, num_tbs, num_sm)

    def hs_golden(self, prim, num_ldst, num_tbs, wg_size, num_sm, num_iters, num_repeats):
        n = self.num_locs_mult(prim, num_tbs, num_sm) * wg_size * (num_ldst + 1)
        g = np.empty(n, dtype=np.float32)
        self.lib.orc_hs_golden(prim, num_ldst, num_tbs, wg_size, num_sm, num_iters, num_repeats,
                               _p(g), n)
        return g

    def bfs(self, row, col, src):
        n = row.size - 1
        out = np.empty(n, dtype=np.int32)
        self.lib.orc_bfs(n, _p(row), _p(col), src, _p(out))
        return out

    def sssp(self, row, col, w, src):
        n = row.size - 1
        out = np.empty(n, dtype=np.int32)
        self.lib.orc_sssp(n, _p(row), _p(col), _p(w), src, _p(out))
        return out

    def pagerank(self, row, col, alpha=0.85, tol=1e-13, max_iter=10000):
        n = row.size - 1
        out = np.empty(n, dtype=np.float64)
        it = self.lib.orc_pagerank_f64(n, _p(row), _p(col), alpha, tol, max_iter, _p(out))
        return out, it

    def pagerank_push_fixpoint(self, row, col, alpha=0.85, tol=1e-13, max_iter=10000):
        n = row.size - 1
        out = np.empty(n, dtype=np.float64)
        it = self.lib.orc_pagerank_push_fixpoint_f64(n, _p(row), _p(col), alpha, tol, max_iter,
                                                     _p(out))
        return out, it

    def gr_write(self, path, row, col, w=None):
        return self.lib.orc_gr_write(path.encode(), row.size - 1, col.size, _p(row), _p(col), _p(w))

    def gr_read(self, path):
        n, m, ety = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = self.lib.orc_gr_read_header(path.encode(), C.byref(n), C.byref(m), C.byref(ety))
        if rc:
            raise IOError(rc)
        row = np.empty(n.value + 1, dtype=np.uint32)
        col = np.empty(m.value, dtype=np.uint32)
        w = np.empty(m.value, dtype=np.uint32) if ety.value else None
        rc = self.lib.orc_gr_read(path.encode(), _p(row), _p(col), _p(w))
        if rc:
            raise IOError(rc)
        return row, col, w

    def par_reduce_i64(self, a):
        t = C.c_int(0)
        s = self.lib.orc_par_reduce_i64(_p(a), a.size, C.byref(t))
        return int(s), t.value

    def par_scan_inclusive_u32(self, a):
        out = np.empty_like(a)
        t = C.c_int(0)
        self.lib.orc_par_scan_inclusive_u32(_p(a), a.size, _p(out), C.byref(t))
        return out, t.value

    def cpu_msbfs(self, row, col, rrow, rcol, srcs):
        """direction-optimizing BFS per source on the host threads: levels [nsrc][n]"""
        n = row.size - 1
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((s.size, n), dtype=np.int32)
        t = C.c_int(0)
        self.lib.orc_cpu_msbfs(n, _p(row), _p(col), _p(rrow), _p(rcol), _p(s), s.size, _p(out), C.byref(t))
        return out, t.value

    def cpu_mssp(self, row, col, w, srcs):
        """Dijkstra per source on the host threads: distances [nsrc][n]"""
        n = row.size - 1
        s = np.ascontiguousarray(srcs, dtype=np.int32)
        out = np.empty((s.size, n), dtype=np.int32)
        t = C.c_int(0)
        self.lib.orc_cpu_mssp(n, _p(row), _p(col), _p(w), _p(s), s.size, _p(out), C.byref(t))
        return out, t.value

    def cpu_pagerank_pull_f32(self, row, rrow, rcol, alpha=0.85, tol=1e-6, max_iter=1000):
        """FP32 pull power iteration on the host threads: (ranks, iterations, threads)"""
        n = row.size - 1
        x = np.empty(n, dtype=np.float32)
        t = C.c_int(0)
        it = self.lib.orc_cpu_pagerank_pull_f32(n, _p(row), _p(rrow), _p(rcol), alpha, tol, max_iter,
                                                _p(x), C.byref(t))
        return x, int(it), t.value
