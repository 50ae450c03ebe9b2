"""Host side of the device serving loop (csrc/serve.hip): submit requests,
poll completions, measure end-to-end latency, and a Poisson load generator."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from .api import _ptr

REDUCE, SCAN, BFS, PR = 0, 1, 2, 3


class Server:
    """A running device serving loop (num_wgs 0: one workgroup per CU).  At the
    co-resident maximum, 2 workgroups per CU, the loop holds every CU's
    registers, and kernels launched while it runs (torch allocations that fill
    memory, generators) wait until it stops: prepare inputs before starting
    it, or use fewer workgroups."""

    def __init__(self, num_wgs: int = 0, ring: int = 1024, budget_units: int = 1 << 20):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.hs_serve_start(num_wgs, ring, budget_units, C.byref(h)), "hs_serve_start")
        self.h = h
        self.submit_t = {}
        self.lat = {}
        self.keep = {}
        self.pending = set()  # submitted, completion not yet seen

    def _submit(self, typ, mem_units, arg, inp, out, n, aux0=0, aux1=0, aux2=0, keep=None, tol=0.0):
        idx = C.c_uint32(0)
        while True:
            if tol != 0:  # a NaN or negative tol reaches the library's check
                rc = self.lib.hs_serve_submit_pagerank(self.h, mem_units, arg, tol, inp, out, n, aux0, aux1,
                                                       aux2, C.byref(idx))
            else:
                rc = self.lib.hs_serve_submit(self.h, typ, mem_units, arg, inp, out, n, aux0, aux1, aux2,
                                              C.byref(idx))
            if rc != L.HS_ERR_NOMEM:
                break
            self.drain(block=False)
            time.sleep(20e-6)
        L.check(rc, "hs_serve_submit")
        self.submit_t[idx.value] = time.perf_counter()
        self.pending.add(idx.value)
        if keep is not None:
            self.keep[idx.value] = keep
        return idx.value

    # requests ------------------------------------------------------------
    def reduce(self, x, out, mem_units=1):
        return self._submit(REDUCE, mem_units, 0, _ptr(x), _ptr(out), x.numel(), keep=(x, out))

    def scan(self, x, out, mem_units=1):
        return self._submit(SCAN, mem_units, 0, _ptr(x), _ptr(out), x.numel(), keep=(x, out))

    def bfs(self, g, src, level, scratch, mem_units=1):
        return self._submit(BFS, mem_units, src, _ptr(g.row), _ptr(level), g.n, _ptr(g.col),
                            _ptr(scratch), keep=(g, level, scratch))

    def pagerank(self, g, iters, rank, scratch, mem_units=1, tol=0.0):
        """`iters` iterations, or fewer when tol > 0: the request completes at
        the first iteration whose L1 change is below tol"""
        return self._submit(PR, mem_units, iters, _ptr(g.rrow), _ptr(rank), g.n, _ptr(g.rcol),
                            _ptr(scratch), _ptr(g.row), keep=(g, rank, scratch), tol=tol)

    # completion ------------------------------------------------------------
    def poll(self, idx):
        st = C.c_uint32(0)
        a, e = C.c_uint64(0), C.c_uint64(0)
        r = self.lib.hs_serve_poll(self.h, idx, C.byref(st), C.byref(a), C.byref(e))
        if r < 0:
            raise L.HSError(-r, "hs_serve_poll")
        if r == 1 and idx in self.pending:
            self.lat[idx] = {"latency_s": time.perf_counter() - self.submit_t[idx],
                             "service_s": (e.value - a.value) / 100e6, "status": st.value}
            self.pending.discard(idx)
            self.keep.pop(idx, None)
        return r == 1

    def drain(self, block=True, timeout=120.0):
        t0 = time.perf_counter()
        while True:
            for i in list(self.pending):
                self.poll(i)
            if not block or not self.pending:
                return
            if time.perf_counter() - t0 > timeout:
                raise TimeoutError("serving loop did not complete the requests")
            time.sleep(20e-6)

    def wait(self, idx, timeout=60.0):
        t0 = time.perf_counter()
        while not self.poll(idx):
            if time.perf_counter() - t0 > timeout:
                raise TimeoutError(f"request {idx}")
            time.sleep(10e-6)
        return self.lat[idx]

    def phases(self, idx):
        """phases a completed request ran (PageRank: iterations + 1; BFS: levels + 1)"""
        ph = C.c_uint32(0)
        L.check(self.lib.hs_serve_phases(self.h, idx, C.byref(ph)), "hs_serve_phases")
        return ph.value

    def stop(self, abort=False):
        if self.h:
            L.check(self.lib.hs_serve_stop(self.h, int(abort)), "hs_serve_stop")
            self.h = None

    def percentiles(self):
        lat = np.array([v["latency_s"] for v in self.lat.values()])
        if lat.size == 0:
            return {}
        svc = np.array([v["service_s"] for v in self.lat.values()])
        return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "n": int(lat.size),
                # admission to completion on the device clock (no host polling, no queueing
                # before admission)
                "service_p50_ms": float(np.percentile(svc, 50) * 1e3),
                "service_p99_ms": float(np.percentile(svc, 99) * 1e3)}


def poisson_arrivals(rate_per_s: float, count: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))
