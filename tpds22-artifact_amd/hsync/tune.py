"""Autotuner for the persistent launches: grid size (workgroups per CU), barrier
kind (fan-in shape) and, for scans, the algorithm and round size.

Each candidate is timed on the caller's actual inputs (the results never
change: every configuration is bit-exact), the fastest is kept in a JSON cache
keyed by device, operation and a size bucket, and the default configuration is
always among the candidates, so the choice is never slower than the default as
measured in the same run.
"""
from __future__ import annotations

import json
import math
import os

from . import api

CACHE_PATH = os.environ.get("HSYNC_TUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "hsync_tune.json"))


def _torch():
    import torch
    return torch


def _time(fn, reps):
    torch = _torch()
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


class Tuner:
    def __init__(self, arena, cache_path: str = CACHE_PATH):
        self.arena = arena
        self.cache_path = cache_path
        try:
            self.cache = json.load(open(cache_path))
        except (OSError, ValueError):
            self.cache = {}
        info = api.device_info()
        self.cus = info["num_cus"]
        self.dev = f"{info['arch']}:{self.cus}"

    def _key(self, op, n):
        return f"{self.dev}|{op}|{int(math.log2(max(n, 1)))}"

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.cache_path), exist_ok=True)
            json.dump(self.cache, open(self.cache_path, "w"), indent=1)
        except OSError:
            pass

    def _pick(self, op, n, candidates, run, reps, refresh):
        key = self._key(op, n)
        if not refresh and key in self.cache:
            return self.cache[key]["config"], self.cache[key]
        results = []
        for cfg in candidates:
            try:
                ms = _time(lambda: run(cfg), reps)
            except api.L.HSError:
                continue
            results.append((ms, cfg))
        results.sort(key=lambda r: r[0])
        default_ms = next(ms for ms, cfg in results if cfg == candidates[0])
        rec = {"config": results[0][1], "ms": results[0][0], "default_ms": default_ms,
               "tried": [{"config": c, "ms": m} for m, c in results]}
        self.cache[key] = rec
        self._save()
        return rec["config"], rec

    def reduce(self, x, refresh=False, reps=20):
        """best (num_wgs, barrier) for hs_reduce_i32 on x; candidates[0] is the default"""
        torch = _torch()
        partials = torch.empty(8 * self.cus, dtype=torch.int32, device=x.device)
        out = torch.empty(1, dtype=torch.int32, device=x.device)
        cands = [{"num_wgs": 0, "barrier": "xcd"}]
        for per_cu in (1, 2, 4):
            for kind in ("xcd", "two_level", "flat"):
                c = {"num_wgs": per_cu * self.cus, "barrier": kind}
                if c not in cands:
                    cands.append(c)

        def run(c):
            api.reduce_i32(self.arena, x, barrier=c["barrier"], num_wgs=c["num_wgs"], out=out,
                           partials=partials, sync=False)
        cfg, rec = self._pick("reduce_i32", x.numel(), cands, run, reps, refresh)
        self.arena.sync()
        return cfg, rec

    def scan(self, x, out, refresh=False, reps=3):
        """best scan configuration: multi-phase barrier scan (grid, barrier, round
        size) or the single-pass look-back scan (grid)"""
        torch = _torch()
        scratch = torch.empty(16 * self.cus, dtype=torch.int32, device=x.device)
        n = x.numel()
        offsets = [0, n]
        cands = [{"algo": "barrier", "num_wgs": 0, "barrier": "xcd",
                  "round_mb": api.get_scan_round_bytes() >> 20}]
        for per_cu in (1, 2, 4):
            for rmb in (32, 64, 192):
                c = {"algo": "barrier", "num_wgs": per_cu * self.cus, "barrier": "xcd", "round_mb": rmb}
                if c not in cands:
                    cands.append(c)
        for per_cu in (2, 4, 8):
            cands.append({"algo": "lookback", "num_wgs": per_cu * self.cus})
        work = {}
        round0 = api.get_scan_round_bytes()

        def run(c):
            if c["algo"] == "barrier":
                api.set_scan_round_bytes(c["round_mb"] << 20)
                api.scan_u32(self.arena, x, inclusive=True, out=out, barrier=c["barrier"],
                             num_wgs=c["num_wgs"], scratch=scratch, sync=False)
            else:
                if "w" not in work:
                    nb = api.C.c_uint64(0)
                    import numpy as np
                    h = np.array(offsets, np.uint64)
                    api.L.check(api.L.load().hs_batch_work_size(h.ctypes.data, 1, api.C.byref(nb)))
                    work["w"] = torch.empty(nb.value, dtype=torch.uint8, device=x.device)
                api.scan_batch_u32(x, offsets, inclusive=True, out=out, num_wgs=c["num_wgs"],
                                   work=work["w"], sync=False)
        try:
            cfg, rec = self._pick("scan_u32", n, cands, run, reps, refresh)
        finally:
            api.set_scan_round_bytes(round0)
        self.arena.sync()
        return cfg, rec
