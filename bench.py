#!/usr/bin/env python3
"""Benchmark of the device-wide synchronization hot path on one MI355X.

Headline (one JSON line on stdout): grid-barrier nanoseconds per episode at
1024 workgroups x 256 threads, 16 global stores per thread per episode, 1e5
episodes per persistent launch, with the XCD-hierarchical barrier.  One step =
one persistent launch.  In the same run: the cooperative-groups grid sync and
the other barrier kinds on the same workload, the workgroup sweep
256/512/1024/2048, the 1e6-episode invariant run, the persistent int32 reduce
(2^24 elements, HBM roofline of the dominant kernel), the persistent inclusive
scan (2^28 elements), and a host-CPU baseline.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--episodes 100000]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tpds22-artifact_amd"))

METRIC = "barrier ns/episode@1024 WGs; reduce GB/s; BFS GTEPS learned vs plain"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_launches(torch, fn, count):
    """average milliseconds of `count` back-to-back launches on the current stream"""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for i in range(count):
        fn(i)
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / count


def cpu_baseline_leg(n_reduce):
    """Host-CPU baseline (oracle library, OpenMP on every host core): the
    barrier episode and the int32 reduce.  A bounded sample of each workload."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import ctypes as C

    import numpy as np
    import oracle_lib
    orc = oracle_lib.load()
    lib = orc.lib
    lib.orc_cpu_barrier_ns.restype = C.c_double
    lib.orc_cpu_barrier_ns.argtypes = [C.c_int, C.POINTER(C.c_int)]
    t = C.c_int(0)
    episodes = 200000
    lib.orc_cpu_barrier_ns(1000, C.byref(t))  # warm-up
    ns = lib.orc_cpu_barrier_ns(episodes, C.byref(t))
    a = orc.gen_uniform_u32(0, n_reduce, 0, 255).astype(np.int32)
    orc.par_reduce_i64(a)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        s, threads = orc.par_reduce_i64(a)
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(ns, 1), "unit": "ns/episode", "cores": t.value, "kind": "port",
            "sample": f"{episodes} episodes of a {t.value}-thread sense-reversing spin barrier, "
                      f"16 stores per thread per episode",
            "reduce_GBps": round(4 * n_reduce / dt / 1e9, 2), "reduce_threads": threads,
            "reduce_sample": f"{reps} sums of 2^24 int32 (OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--episodes", type=int, default=100000)
    ap.add_argument("--no-extras", action="store_true", help="headline barrier only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the synchronization "
              "primitives are per-device and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch
    import hsync as hs
    hs.load()
    info = hs.device_info()
    log("device", info)
    cus = info["num_cus"]
    arena = hs.Arena(max_blocks=2048, num_groups=cus)
    extra = {"device": info}

    # ---------------- headline: barrier episode at 1024 WGs x 256 threads
    W, WG, L, E = 1024, 256, 16, args.episodes
    storage = torch.zeros(2048 * WG * (L + 1), dtype=torch.float32, device="cuda")

    def bar_launch(kind, wgs=W, eps=E, check=False, v=None):
        return hs.barrier_bench(arena, kind, storage, wgs, WG, L, eps, check=check, violations=v,
                                sync=False)

    def measure_barrier(kind, wgs=W, eps=E, steps=args.steps, warmup=args.warmup):
        for _ in range(warmup):
            bar_launch(kind, wgs, eps)
        arena.sync()
        torch.cuda.synchronize()
        ms = timed_launches(torch, lambda i: bar_launch(kind, wgs, eps), steps)
        torch.cuda.synchronize()
        arena.sync()
        return ms, ms * 1e6 / eps

    log("headline: xcd barrier", W, "WGs")
    ms_step, ns_xcd = measure_barrier("xcd")
    log(f"  xcd: {ns_xcd:.1f} ns/episode ({ms_step:.2f} ms/step)")
    results = {"xcd": ns_xcd}
    if not args.no_extras:
        for kind in ("coop", "two_level", "flat", "xcd_flat_release", "lf", "tree"):
            eps = E if kind in ("coop", "two_level", "flat", "xcd_flat_release") else E // 10
            _, ns = measure_barrier(kind, eps=eps, steps=2, warmup=1)
            results[kind] = ns
            log(f"  {kind}: {ns:.1f} ns/episode")
        extra["barrier_ns_per_episode_1024"] = {k: round(v, 1) for k, v in results.items()}
        extra["xcd_speedup_vs_coop_1024"] = round(results["coop"] / ns_xcd, 3)
        sweep = {}
        for wgs in (256, 512, 1024, 2048):
            row = {}
            for kind in ("xcd", "coop", "two_level", "flat"):
                try:
                    _, ns = measure_barrier(kind, wgs=wgs, eps=20000, steps=1, warmup=1)
                    row[kind] = round(ns, 1)
                except hs.HSError as ex:
                    row[kind] = f"refused: {ex}"
            sweep[str(wgs)] = row
            log("  sweep", wgs, row)
        extra["barrier_sweep_ns_per_episode"] = sweep
        # invariant run: 1e6 episodes at 1024 WGs, every episode checked
        v = torch.zeros(hs.check_words(arena.num_groups), dtype=torch.int32, device="cuda")
        t0 = time.perf_counter()
        bar_launch("xcd", eps=1000000, check=True, v=v)
        arena.sync()
        extra["invariant_run"] = {"barrier": "xcd", "wgs": W, "episodes": 1000000,
                                  "early_passes": int(v[0].item()),
                                  "seconds": round(time.perf_counter() - t0, 2)}
        log("  invariants", extra["invariant_run"])

    # ---------------- reduce: 2^24 int32 uniform [0,255] (dominant HBM kernel)
    N = 1 << 24
    nbuf = 8  # 8 x 64 MiB = 512 MiB rotated > 256 MiB Infinity Cache: HBM-cold reads
    bufs = [torch.empty(N, dtype=torch.int32, device="cuda") for _ in range(nbuf)]
    for i, b in enumerate(bufs):
        hs.gen_uniform_u32(b, i, 0, 255)
    partials = torch.empty(4096, dtype=torch.int32, device="cuda")
    out = torch.empty(1, dtype=torch.int32, device="cuda")

    def red(i, warm=False):
        hs.reduce_i32(arena, bufs[0 if warm else i % nbuf], barrier="xcd", out=out,
                      partials=partials, sync=False)

    for i in range(20):
        red(i)
    arena.sync()
    R = 200
    ms_cold = timed_launches(torch, red, R)
    ms_warm = timed_launches(torch, lambda i: red(i, True), R)
    arena.sync()
    # correctness of the timed kernel on buffer 0
    s_gpu = hs.reduce_i32(arena, bufs[0])
    s_ref = int(bufs[0].sum(dtype=torch.int64).item())
    assert s_gpu == s_ref, (s_gpu, s_ref)
    gbps_cold = 4 * N / (ms_cold * 1e-3) / 1e9
    gbps_warm = 4 * N / (ms_warm * 1e-3) / 1e9
    log(f"reduce 2^24: cold {ms_cold*1e3:.2f} us {gbps_cold:.0f} GB/s, warm {ms_warm*1e3:.2f} us "
        f"{gbps_warm:.0f} GB/s")
    extra["reduce_2p24"] = {"GBps_hbm_cold": round(gbps_cold, 1), "GBps_mall_warm": round(gbps_warm, 1),
                            "us_per_launch_cold": round(ms_cold * 1e3, 2),
                            "us_per_launch_warm": round(ms_warm * 1e3, 2), "grid": "512 WGs x 256",
                            "barrier": "xcd", "sum_matches": True}
    del bufs

    # ---------------- scan: 2^28 int32 uniform [0,15], inclusive
    if not args.no_extras:
        NS = 1 << 28
        x = torch.empty(NS, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        hs.gen_uniform_u32(x, 0, 0, 15)
        scratch = torch.empty(8192, dtype=torch.int32, device="cuda")
        for _ in range(2):
            hs.scan_u32(arena, x, inclusive=True, out=y, scratch=scratch, sync=False)
        arena.sync()
        ms_scan = timed_launches(torch, lambda i: hs.scan_u32(arena, x, inclusive=True, out=y,
                                                              scratch=scratch, sync=False), 5)
        arena.sync()
        ok = int(y[-1].item()) == int(x.sum(dtype=torch.int64).item())
        gb_scan = 8 * NS / (ms_scan * 1e-3) / 1e9
        extra["scan_2p28_inclusive"] = {"GBps": round(gb_scan, 1), "ms": round(ms_scan, 3),
                                        "roofline_frac": round(gb_scan / HBM_PEAK_GBPS, 3),
                                        "last_equals_total": ok}
        log("scan 2^28:", extra["scan_2p28_inclusive"])
        del x, y

    # ---------------- CPU baseline (host cores, oracle library)
    cpu = cpu_baseline_leg(N)
    log("cpu baseline", cpu)

    line = {
        "metric": METRIC,
        "value": round(ns_xcd, 1),
        "unit": "ns/episode",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"barrier microbench: {W} WGs x {WG} threads, {L} global stores per "
                               f"thread per episode, {E} episodes per persistent launch, "
                               "XCD-hierarchical barrier",
                   "wgs": W, "wg_size": WG, "stores_per_thread": L, "episodes": E},
        "roofline": {"kernel": "k_reduce_i32 (2^24 int32, HBM-cold rotation)", "bound": "hbm",
                     "achieved": round(gbps_cold, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps_cold / HBM_PEAK_GBPS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    arena.close()


if __name__ == "__main__":
    main()
