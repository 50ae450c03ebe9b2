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


def profiled_traffic():
    """HBM bytes per k_reduce_i32 launch from the committed rocprofv3 counter
    runs (FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy), or None"""
    import re
    path = os.path.join(REPO, "profiles", "k_reduce_i32_2p24_hbm_cold.md")
    try:
        m = re.search(r"Reduce 2\^24 traffic per launch = .*= ([0-9.]+) MB", open(path).read())
    except OSError:
        return None
    return int(float(m.group(1)) * 1e6) if m else None


def warm_clocks(torch, fn, seconds=0.3):
    """run fn() repeatedly for `seconds`: the GPU's clocks ramp up under sustained
    load, and a timed region that starts cold reads slow"""
    t0 = time.perf_counter()
    i = 0
    while time.perf_counter() - t0 < seconds:
        fn(i)
        i += 1
        if i % 8 == 0:
            torch.cuda.current_stream().synchronize()
    torch.cuda.current_stream().synchronize()


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


def _median_seconds(fn, reps):
    ts = []
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2], out


def cpu_scan_leg(x_host, y_gpu_host):
    """Host baseline of the 2^28 inclusive scan (OpenMP two-pass blocked scan on
    every host core), median of 5 after a warm-up; its output is compared with
    the GPU's."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np
    import oracle_lib
    orc = oracle_lib.load()
    u = x_host.view(np.uint32)
    orc.par_scan_inclusive_u32(u)
    dt, (out, threads) = _median_seconds(lambda: orc.par_scan_inclusive_u32(u), 5)
    return {"GBps": round(8 * u.size / dt / 1e9, 2), "ms": round(dt * 1e3, 1), "threads": threads,
            "sample": "inclusive scan of the same 2^28 int32 input, median of 5 after a warm-up",
            "equals_gpu": bool(np.array_equal(out, y_gpu_host.view(np.uint32)))}


def cpu_graph_leg(g, w, srcs, levels_gpu, dist_gpu, rank_gpu, bfs_edges, sssp_edges_16):
    """Host baselines of the graph workloads on every host core: direction-
    optimizing BFS (64 sources, one per thread at a time), binary-heap Dijkstra
    (a sample of 16 of the sources), FP32 pull PageRank to the same L1 tolerance.
    Each result is compared with the GPU's on the full graph."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np
    import oracle_lib
    orc = oracle_lib.load()
    out = {}
    orc.cpu_msbfs(g.h_row, g.h_col, g.h_rrow, g.h_rcol, srcs[:4])  # warm-up
    dt, (lv, threads) = _median_seconds(
        lambda: orc.cpu_msbfs(g.h_row, g.h_col, g.h_rrow, g.h_rcol, srcs), 5)
    out["bfs_64"] = {"ms": round(dt * 1e3, 1), "GTEPS": round(bfs_edges / dt / 1e9, 2), "threads": threads,
                     "sample": "64 sources, direction-optimizing BFS per source, median of 5 after a warm-up",
                     "levels_equal_gpu": bool(np.array_equal(lv, levels_gpu))}
    del lv
    log("  cpu bfs", out["bfs_64"])
    dt, (d, threads) = _median_seconds(lambda: orc.cpu_mssp(g.h_row, g.h_col, w, srcs[:16]), 1)
    out["sssp_16"] = {"ms": round(dt * 1e3, 1), "GTEPS": round(sssp_edges_16 / dt / 1e9, 3), "threads": threads,
                      "sample": "16 of the 64 sources, binary-heap Dijkstra per source, one run",
                      "dist_equal_gpu": bool(np.array_equal(d, dist_gpu))}
    del d
    log("  cpu sssp", out["sssp_16"])
    orc.cpu_pagerank_pull_f32(g.h_row, g.h_rrow, g.h_rcol, 0.85, 1e-6, 2)  # warm-up
    dt, (x, iters, threads) = _median_seconds(
        lambda: orc.cpu_pagerank_pull_f32(g.h_row, g.h_rrow, g.h_rcol, 0.85, 1e-6, 1000), 5)
    # the GPU's FP32 ranks against the FP64 oracle iterated to an L1 change < 1e-12
    ref64, it64 = orc.pagerank(g.h_row, g.h_col, alpha=0.85, tol=1e-12)
    out["pagerank"] = {"ms": round(dt * 1e3, 1), "iterations": iters, "threads": threads,
                       "sample": "FP32 pull power iteration to L1 change < 1e-6, median of 5 after a warm-up",
                       "max_abs_diff_vs_gpu": float(np.abs(x.astype(np.float64) - rank_gpu).max()),
                       "gpu_max_abs_err_vs_fp64_oracle": float(np.abs(rank_gpu - ref64).max()),
                       "fp64_oracle_iterations": int(it64)}
    log("  cpu pagerank", out["pagerank"])
    return out


def bench_batched_scan(torch, hs):
    """512 inclusive scans of log-uniform lengths 2^10..2^22 in one work-stealing
    launch (8 bytes per element of algorithmic traffic)."""
    import numpy as np
    rng = np.random.default_rng(0)
    lens = np.exp2(rng.uniform(10, 22, 512)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    hs.gen_uniform_u32(x, 0, 0, 15)
    import ctypes as C
    lib = hs.load()
    nb = C.c_uint64(0)
    lib.hs_batch_work_size(off.ctypes.data, 512, C.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    # the autotuned launch: the grid hs_tune_scan_batch_u32 picks on this batch
    tune = hs.tune_scan_batch_u32(x, off, y, inclusive=True, reps=5)
    wgs = tune["num_wgs"]
    for _ in range(2):
        hs.scan_batch_u32(x, off, out=y, work=work, num_wgs=wgs)
    ms = timed_launches(torch, lambda i: hs.scan_batch_u32(x, off, out=y, work=work, num_wgs=wgs,
                                                           sync=False), 10)
    hs.scan_batch_u32(x, off, out=y, work=work, num_wgs=wgs)  # status check
    ends = torch.as_tensor(off[1:].astype(np.int64) - 1, device="cuda")
    starts = torch.as_tensor(off[:-1].astype(np.int64), device="cuda")
    csum = torch.cumsum(x.to(torch.int64), 0)
    seg = csum[ends] - torch.where(starts > 0, csum[(starts - 1).clamp(min=0)], torch.zeros_like(starts))
    ok = bool(torch.equal(y[ends].to(torch.int64), seg))
    return {"elements": n, "ms": round(ms, 3), "GBps": round(8 * n / (ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(8 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3), "segment_ends_ok": ok,
            "tuned_grid": wgs, "tuner_ms": round(tune["ms"], 3), "tuner_default_ms": round(tune["default_ms"], 3)}


# R-MAT (a, b, c, edge factor): the families the schedule network trains on and
# one it never sees
SCHED_TRAIN_FAMILIES = ((0.57, 0.19, 0.19, 16), (0.57, 0.19, 0.19, 4), (0.45, 0.22, 0.22, 16),
                        (0.25, 0.25, 0.25, 16))
SCHED_HELD_OUT = ((0.57, 0.19, 0.19, 48), (0.65, 0.15, 0.15, 16), (0.75, 0.10, 0.10, 8))
# families that only the judge's margins are chosen on (neither trained on nor tested)
SCHED_VAL_FAMILIES = ((0.60, 0.17, 0.17, 32), (0.70, 0.12, 0.12, 12))


def _rmat_sources(G, np, row, seed):
    deg = np.diff(row.astype(np.int64))
    return [int(v) for v in np.random.default_rng(seed).choice(np.nonzero(deg > 0)[0], 64, replace=False)]


def bench_bfs(torch, hs, arena, scale, cpu_out=None):
    """64-source BFS on R-MAT (a,b,c)=(0.57,0.19,0.19), edge factor 16: plain
    heuristic schedule vs the learned schedule (MLP trained in this run by the HIP
    training kernels on the per-action level times of smaller R-MAT graphs of four
    families, judge margins chosen per size band on two validation families) vs the per-level
    best; the same comparison on graphs of three families left out of training."""
    import numpy as np
    import hsync.graph as G
    import hsync.sched as S
    t0 = time.perf_counter()
    # training graphs of the training families, and validation graphs of two
    # other families in a small and a large size band: the judge takes a margin
    # per band, each chosen on its band's graphs of families the network never saw
    train, val_small, val_large = [], [], []  # validation: one sample list per graph
    train_scales, small_scale, large_scale = (scale - 7, scale - 6, scale - 5, scale - 3, scale - 2), scale - 4, scale - 1

    def samples_of(fam, sc, seed):
        a, b, c, ef = fam
        row, col = G.rmat_csr(sc, ef, seed, a=a, b=b, c=c)
        gt = G.Graph(row, col)
        smp = S.collect(arena, gt, _rmat_sources(G, np, row, seed))
        del gt
        return smp

    for k, fam in enumerate(SCHED_TRAIN_FAMILIES):
        for sc in train_scales:
            train.extend(samples_of(fam, sc, 100 * sc + k))
    for k, fam in enumerate(SCHED_VAL_FAMILIES):
        val_small.append(samples_of(fam, small_scale, 700 + k))
        val_large.append(samples_of(fam, large_scale, 800 + k))
    X, C, _ = S.to_arrays(train)
    mlp = S.MLP(seed=5)
    hist = mlp.fit_cost(X, C, epochs=1500, lr=2e-3)
    threshold = (small_scale + large_scale) / 2  # in log2(n + 1)
    m_small, r_small, m_large, r_large = mlp.choose_judge(val_small, val_large, threshold)
    t_train = time.perf_counter() - t0
    fin = lambda v: v if v != float("inf") else "inf"  # noqa: E731
    res = {"scale": scale, "sources": 64, "train_samples": len(train),
           "train_scales": list(train_scales), "val_scales": {"small": small_scale, "large": large_scale},
           "val_families": [list(f) for f in SCHED_VAL_FAMILIES],
           "train_loss_hist": [round(v, 4) for v in hist[::5]], "train_seconds": round(t_train, 1),
           "judge": {"margin_small": fin(m_small), "margin_large": fin(m_large), "threshold_log2n": threshold,
                     "val_small_plain_over_learned": round(1.0 / r_small, 3),
                     "val_large_plain_over_learned": round(1.0 / r_large, 3)}}

    def compare(g, srcs, levels, work, reps=5):
        """ms and GTEPS of the plain, learned, push-only, pull-only and per-level
        best schedules; every run's levels equal to the plain schedule's"""
        G.msbfs(arena, g, srcs, schedule="heuristic", levels=levels, work=work)
        ref = levels.clone()
        edges = G.reached_edges(ref, torch.as_tensor(g.out_degree.astype(np.float64), device="cuda"))
        smp = S.collect(arena, g, srcs)
        best = [int(np.argmin(s["t"])) for s in smp]
        out = {"traversed_edges": edges}
        d_srcs = G.device_sources(srcs)  # timed launches read the sources from the device
        for name, sched, forced in (("plain", "heuristic", None), ("learned", "learned", None),
                                    ("push_only", "push", None), ("pull_only", "pull", None),
                                    ("per_level_best", "forced", best)):
            G.msbfs(arena, g, srcs, schedule=sched, mlp=mlp.w, forced=forced, levels=levels, work=work)
            ms = timed_launches(torch, lambda i: G.msbfs(arena, g, d_srcs, schedule=sched, mlp=mlp.w,
                                                         forced=forced, levels=levels, work=work, sync=False), reps)
            arena.sync()
            out[name] = {"ms": round(ms, 3), "GTEPS": round(edges / (ms * 1e-3) / 1e9, 2),
                         "levels_identical": bool(torch.equal(levels, ref))}
        out["learned_vs_plain"] = round(out["plain"]["ms"] / out["learned"]["ms"], 3)
        out["best_actions"] = best
        return out, ref

    # graphs of the three held-out families, four scales below the headline graph
    nb = hs.api.C.c_uint64(0)
    held, tp, tl = [], 0.0, 0.0
    for k, (a, b, c, ef) in enumerate(SCHED_HELD_OUT):
        row, col = G.rmat_csr(scale - 4, ef, 300 + k, a=a, b=b, c=c)
        gh = G.Graph(row, col)
        lv_h = torch.empty((64, gh.n), dtype=torch.int32, device="cuda")
        hs.load().hs_msbfs_work_size(gh.n, gh.m, 64, hs.api.C.byref(nb))
        wk_h = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
        r, _ = compare(gh, _rmat_sources(G, np, row, 300 + k), lv_h, wk_h, reps=10)
        r.update({"rmat_abc": [a, b, c], "edge_factor": ef, "scale": scale - 4})
        r.pop("best_actions")
        held.append(r)
        tp += r["plain"]["ms"]
        tl += r["learned"]["ms"]
        del gh, lv_h, wk_h
    res["held_out_families"] = held
    res["held_out_learned_vs_plain"] = round(tp / tl, 3)
    log("  learned schedule on held-out families:", res["held_out_learned_vs_plain"], "judge", res["judge"])
    # the headline graph
    g = G.Graph.rmat(scale, 16, 1)
    deg = g.out_degree
    srcs = [int(v) for v in np.random.default_rng(7).choice(np.nonzero(deg > 0)[0], 64, replace=False)]
    outdeg_t = torch.as_tensor(deg.astype(np.float64), device="cuda")
    levels = torch.empty((64, g.n), dtype=torch.int32, device="cuda")
    hs.load().hs_msbfs_work_size(g.n, g.m, 64, hs.api.C.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    res.update({"nnodes": g.n, "nedges": g.m})
    cmp, ref = compare(g, srcs, levels, work)
    edges = cmp.pop("traversed_edges")
    best = cmp.pop("best_actions")
    res["traversed_edges"] = edges
    res.update(cmp)
    # the plain persistent BFS (LonestarGPU-style: one source per launch, push
    # only, one grid barrier per level) over the same 64 sources
    lvl = torch.empty(g.n, dtype=torch.int32, device="cuda")
    nbw = hs.api.C.c_uint64(0)
    hs.load().hs_bfs_work_size(g.n, g.m, hs.api.C.byref(nbw))
    wk = torch.empty(nbw.value, dtype=torch.uint8, device="cuda")
    G.bfs(arena, g, srcs[0], level=lvl, work=wk)
    ms_pp = timed_launches(torch, lambda i: G.bfs(arena, g, srcs[i], level=lvl, work=wk, sync=False), 64) * 64
    arena.sync()
    same_pp = True
    for i in (0, 21, 42, 63):
        G.bfs(arena, g, srcs[i], level=lvl, work=wk)
        same_pp &= bool(torch.equal(lvl, ref[i]))
    res["plain_persistent_64x1"] = {"ms": round(ms_pp, 3), "GTEPS": round(edges / (ms_pp * 1e-3) / 1e9, 2),
                                    "levels_identical_sampled": same_pp}
    res["learned_batched_vs_plain_persistent"] = round(ms_pp / res["learned"]["ms"], 3)
    # SSSP on the same graph with weights uniform [1,255]: the 64 sources in one
    # batched launch vs the plain persistent near/far SSSP one source at a time
    gw = G.Graph(g.h_row, g.h_col, G.weights(g.m, 1001), transposed=False)
    dist = torch.empty((64, g.n), dtype=torch.int32, device="cuda")
    nbm = hs.api.C.c_uint64(0)
    hs.load().hs_mssp_work_size(g.n, g.m, 64, hs.api.C.byref(nbm))
    wm = torch.empty(nbm.value, dtype=torch.uint8, device="cuda")
    st = torch.zeros(4, dtype=torch.int32, device="cuda")
    G.mssp(arena, gw, srcs, dist=dist, work=wm, stats=st)
    d_srcs = G.device_sources(srcs)
    ms_ms = timed_launches(torch, lambda i: G.mssp(arena, gw, d_srcs, dist=dist, work=wm, stats=st, sync=False), 3)
    arena.sync()
    e_sssp = G.reached_edges(dist, outdeg_t)
    d1 = torch.empty(g.n, dtype=torch.int32, device="cuda")
    nbs = hs.api.C.c_uint64(0)
    hs.load().hs_sssp_work_size(g.n, g.m, hs.api.C.byref(nbs))
    ws_ = torch.empty(nbs.value, dtype=torch.uint8, device="cuda")
    G.sssp(arena, gw, srcs[0], dist=d1, work=ws_)
    ms_one = timed_launches(torch, lambda i: G.sssp(arena, gw, srcs[i], dist=d1, work=ws_, sync=False), 8)
    arena.sync()
    same_ss = True
    for i in (0, 31, 63):
        G.sssp(arena, gw, srcs[i], dist=d1, work=ws_)
        same_ss &= bool(torch.equal(d1, dist[i]))
    res["sssp_batched_64"] = {"ms": round(ms_ms, 3), "GTEPS": round(e_sssp / (ms_ms * 1e-3) / 1e9, 2),
                              "rounds": int(st[0].item()), "traversed_edges": e_sssp}
    res["sssp_plain_persistent_64x1"] = {"ms_per_source": round(ms_one, 3), "ms_64": round(64 * ms_one, 2),
                                         "GTEPS": round(e_sssp / (64 * ms_one * 1e-3) / 1e9, 2),
                                         "rows_identical_sampled": same_ss}
    res["sssp_batched_vs_plain_persistent"] = round(64 * ms_one / ms_ms, 2)
    # PageRank (FP32, alpha 0.85, until the L1 change < 1e-6, normalized) on the
    # same graph: algorithmic bytes per iteration 8 per in-edge (rcol + gathered
    # contribution) + 20 per vertex (rrow, row, x in, x out, contribution out)
    nbp = hs.api.C.c_uint64(0)
    hs.load().hs_pagerank_work_size(g.n, g.m, hs.api.C.byref(nbp))
    wp = torch.empty(nbp.value, dtype=torch.uint8, device="cuda")
    rank = torch.empty(g.n, dtype=torch.float32, device="cuda")
    _, pst = G.pagerank(arena, g, rank=rank, work=wp)
    ms_pr = timed_launches(torch, lambda i: G.pagerank(arena, g, rank=rank, work=wp, sync=False), 3)
    arena.sync()
    _, pst = G.pagerank(arena, g, rank=rank, work=wp)
    pr_it = int(pst[0].item())
    pr_bytes = (8 * g.m + 20 * g.n) * pr_it
    res["pagerank"] = {"ms": round(ms_pr, 3), "iterations": pr_it, "ms_per_iteration": round(ms_pr / pr_it, 4),
                       "GBps": round(pr_bytes / (ms_pr * 1e-3) / 1e9, 1),
                       "roofline_frac": round(pr_bytes / (ms_pr * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3),
                       "rank_sum": round(float(rank.double().sum().item()), 6)}
    log("  pagerank", res["pagerank"])
    if cpu_out is not None:
        e16 = G.reached_edges(dist[:16], outdeg_t)
        cpu_out.update(cpu_graph_leg(g, gw.h_w, srcs, ref.cpu().numpy(),
                                     dist[:16].cpu().numpy(), rank.cpu().numpy().astype(np.float64),
                                     edges, e16))
        cpu_out["bfs_64"]["gpu_learned_batched_vs_cpu"] = round(cpu_out["bfs_64"]["ms"] / res["learned"]["ms"], 1)
        cpu_out["sssp_16"]["gpu_batched_vs_cpu_per_source"] = round(
            (cpu_out["sssp_16"]["ms"] / 16) / (ms_ms / 64), 1)
        cpu_out["pagerank"]["gpu_vs_cpu"] = round(cpu_out["pagerank"]["ms"] / ms_pr, 1)
    del gw, dist, wm, rank, wp
    tr = G.new_trace()
    G.msbfs(arena, g, srcs, schedule="learned", mlp=mlp.w, trace=tr, levels=levels, work=work)
    res["learned_actions"] = [r["act"] for r in G.decode_trace(tr)]
    tr = G.new_trace()
    G.msbfs(arena, g, srcs, schedule="heuristic", trace=tr, levels=levels, work=work)
    res["plain_actions"] = [r["act"] for r in G.decode_trace(tr)]
    res["best_actions"] = best
    del g
    return res


def cpu_serving_leg(entries, graphs):
    """Host baseline of the serving mix, one pool entry after another on every
    host core: the result each entry's requests must produce (reduce and
    inclusive scan of the entry's generated input, BFS levels from vertex 0,
    FP32 pull PageRank to L1 change < 1e-6).  Returns the results and the host
    time of one pass over the pool."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np
    import oracle_lib
    orc = oracle_lib.load()
    t0 = time.perf_counter()
    per_graph = {}
    exp = []
    threads = 1
    for k, key in entries:
        if k < 2:
            seed, n = key
            h = orc.gen_uniform_u32(seed, n, 0, 255)
            if k == 0:
                s_, threads = orc.par_reduce_i64(h.astype(np.int32))
                exp.append(np.array([s_ & 0xFFFFFFFF], np.uint32).view(np.int32))
            else:
                out, threads = orc.par_scan_inclusive_u32(h)
                exp.append(out.view(np.int32))
        else:
            gi = key
            if (k, gi) not in per_graph:
                g = graphs[gi]
                if k == 2:
                    lv, threads = orc.cpu_msbfs(g.h_row, g.h_col, g.h_rrow, g.h_rcol, [0])
                    per_graph[(k, gi)] = lv[0]
                else:
                    x, _, threads = orc.cpu_pagerank_pull_f32(g.h_row, g.h_rrow, g.h_rcol, 0.85, 1e-6, 1000)
                    per_graph[(k, gi)] = x
            exp.append(per_graph[(k, gi)])
    return exp, time.perf_counter() - t0, threads


def bench_serving(torch, hs, seconds=4.0, rate=500.0, num_wgs=512):
    """Poisson stream of mixed requests into the device serving loop: int32
    reduce / scan of 2^12..2^26 elements (log-uniform), BFS and FP32 PageRank
    (damping 0.85, until the L1 change of an iteration is below 1e-6, at most
    1000 iterations) on R-MAT graphs of scale 12..20, admitted by memory
    (budget units of 1 MiB); end-to-end latency percentiles and throughput.
    p50_ms / p99_ms count from each request's submission; the *_from_arrival
    figures count from its scheduled Poisson arrival, so they also hold the
    load generator's own delays (host stalls; a request waiting for the
    previous request of its pool entry, whose scratch it shares -- concurrent
    PageRank iterations on one scratch would never reach the tolerance).
    Every request writes an output of its own, and after the stream every
    output is compared with the host leg's result for its entry: integers
    bit for bit, PageRank within 1e-6 max-abs."""
    import numpy as np
    import hsync.graph as G
    from hsync.serve import Server, poisson_arrivals
    rng = np.random.default_rng(3)
    graphs = [G.Graph.rmat(sc, 16, 50 + sc) for sc in range(12, 21)]
    pool = []
    entries = []  # what each entry computes, for the host leg
    for i in range(144):
        k = i % 4
        if k < 2:
            n = int(2 ** rng.uniform(12, 26))
            x = torch.empty(n, dtype=torch.int32, device="cuda")
            hs.gen_uniform_u32(x, i, 0, 255)
            pool.append((k, x, (1 if k == 0 else n, torch.int32), None, max(1, ((8 if k else 4) * n) >> 20)))
            entries.append((k, (i, n)))
        else:
            gi = (i // 4) % len(graphs)
            g = graphs[gi]
            dt = torch.int32 if k == 2 else torch.float32
            scratch = torch.zeros(2 * g.n, dtype=dt, device="cuda")
            pool.append((k, g, (g.n, dt), scratch, max(1, (12 * g.n) >> 20)))
            entries.append((k, gi))
    order = rng.permutation(len(pool))
    arrivals = poisson_arrivals(rate, int(rate * seconds), seed=5)
    # one output per request (about 15 GB at these sizes), allocated before the loop starts
    outs = []
    for j in range(len(arrivals)):
        size, dt = pool[int(order[j % len(pool)])][2]
        outs.append(torch.zeros(size, dtype=dt, device="cuda"))
    torch.cuda.synchronize()
    srv = Server(num_wgs=num_wgs, ring=1024, budget_units=4096)
    last = {}  # pool entry -> its latest request
    sched = {}  # request -> scheduled arrival
    req_of = []  # arrival j -> (request, entry)
    t0 = time.perf_counter()
    try:
        for j, at in enumerate(arrivals):
            while time.perf_counter() - t0 < at:
                srv.drain(block=False)
            e = int(order[j % len(pool)])
            k, a, _, scratch, mu = pool[e]
            out = outs[j]
            t_wait = time.perf_counter()
            while e in last and last[e] in srv.pending:  # polls every request meanwhile
                srv.drain(block=False)
                if time.perf_counter() - t_wait > 60:
                    raise TimeoutError(f"request {last[e]}")
            if k == 0:
                idx = srv.reduce(a, out, mem_units=mu)
            elif k == 1:
                idx = srv.scan(a, out, mem_units=mu)
            elif k == 2:
                idx = srv.bfs(a, 0, out, scratch, mem_units=mu)
            else:
                idx = srv.pagerank(a, 1000, out, scratch, mem_units=mu, tol=1e-6)
            sched[idx] = t0 + at
            last[e] = idx
            req_of.append((idx, e))
        srv.drain(timeout=120)
        wall = time.perf_counter() - t0
        p = srv.percentiles()
        lag = np.array([srv.submit_t[i] - sched[i] for i in sched])
        arr = np.array([srv.lat[i]["latency_s"] + srv.submit_t[i] - sched[i] for i in sched])
        p.update({"p50_ms_from_arrival": float(np.percentile(arr, 50) * 1e3),
                  "p99_ms_from_arrival": float(np.percentile(arr, 99) * 1e3),
                  "max_submit_lag_ms": float(lag.max() * 1e3)})
        p.update({"offered_rate_per_s": rate, "throughput_req_per_s": round(len(arrivals) / wall, 1),
                  "statuses_ok": all(v["status"] == 0 for v in srv.lat.values()),
                  "mix": "reduce/scan 2^12..2^26 int32, BFS and PageRank (to 1e-6 L1) on R-MAT 12..20"})
    finally:
        srv.stop()
    # every request's output against its entry's host result
    exp, cpu_s, threads = cpu_serving_leg(entries, graphs)
    bad, pr_err = 0, 0.0
    exp_dev = {}
    for j, (idx, e) in enumerate(req_of):
        k = pool[e][0]
        if e not in exp_dev:
            exp_dev[e] = torch.from_numpy(exp[e]).to("cuda")
        if k == 3:
            err = float((outs[j].double() - exp_dev[e].double()).abs().max().item())
            pr_err = max(pr_err, err)
            bad += err > 1e-6
        else:
            bad += not torch.equal(outs[j], exp_dev[e])
    p.update({"requests_checked": len(req_of), "results_wrong": int(bad),
              "pagerank_max_abs_vs_host_fp32": pr_err,
              "cpu_one_pass_over_pool_ms": round(cpu_s * 1e3, 1), "cpu_threads": threads})
    del outs, exp_dev
    return {k2: (round(v, 3) if isinstance(v, float) and k2 != "pagerank_max_abs_vs_host_fp32" else v)
            for k2, v in p.items()}


# (barrier kind, stores per thread, workgroups) of `--profile barrier_counters`,
# in launch order
BARRIER_COUNTER_RUNS = [("xcd", 0, 1024), ("xcd", 16, 1024), ("flat", 0, 1024), ("flat", 16, 1024),
                        ("two_level", 0, 1024), ("xcd", 0, 2048), ("xcd", 16, 2048)]


def profile_only(torch, hs, what):
    """The kernels of one workload and nothing else, for rocprofv3 (kernel trace
    or counters).  Every family also runs one plain 64 MiB device copy
    (torch's copy kernel): its FETCH_SIZE / WRITE_SIZE readings against the known
    64 MiB calibrate the counters."""
    arena = hs.Arena(max_blocks=2048)
    n = 1 << 24
    a = torch.empty(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    hs.gen_uniform_u32(a, 0, 0, 255)
    b.copy_(a)  # calibration: 64 MiB read + 64 MiB written
    if what == "reduce":
        # the bench's timed region: warm clocks, then 20 replays of a hipGraph of
        # 64 launches rotating over 8 x 64 MiB buffers (HBM-cold), default grid
        from hsync.graphcap import Replay
        bufs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(8)]
        for i, x in enumerate(bufs):
            hs.gen_uniform_u32(x, i, 0, 255)
        partials = torch.empty(4096, dtype=torch.int32, device="cuda")
        out = torch.empty(1, dtype=torch.int32, device="cuda")
        g = Replay(lambda: [hs.reduce_i32(arena, bufs[i % 8], barrier="xcd", out=out, partials=partials,
                                          sync=False) for i in range(64)])
        warm_clocks(torch, lambda i: g.replay())
        for _ in range(20):
            g.replay()
        arena.sync()
        del g
        # practical ceilings with counters: a 1 GiB device copy and a 1 GiB read
        x = torch.empty(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        hs.gen_uniform_u32(x, 0, 0, 255)
        for _ in range(5):
            y.copy_(x)
        for _ in range(5):
            hs.reduce_i32(arena, x, out=out, partials=partials, sync=False)
    elif what == "scan":
        import numpy as np
        x = torch.empty(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        hs.gen_uniform_u32(x, 0, 0, 15)
        for _ in range(4):
            hs.scan_u32(arena, x, inclusive=True, out=y, sync=False)
        o1 = np.array([0, 1 << 28], dtype=np.uint64)
        for _ in range(4):
            hs.scan_batch_u32(x, o1, out=y, sync=False)
    elif what in ("pagerank", "graph"):
        import numpy as np
        import hsync.graph as G
        g = G.Graph.rmat(22, 16, 1)
        if what == "pagerank":
            for _ in range(3):
                G.pagerank(arena, g, alpha=0.85, tol=1e-6, sync=False)
        else:
            srcs = [int(v) for v in np.random.default_rng(7).choice(np.nonzero(g.out_degree > 0)[0], 8,
                                                                    replace=False)]
            gw = G.Graph(g.h_row, g.h_col, G.weights(g.m, 1001), transposed=False)
            for s_ in srcs:
                G.bfs(arena, g, s_, sync=False)
            for s_ in srcs[:4]:
                G.sssp(arena, gw, s_, sync=False)
            srcs64 = [int(v) for v in np.random.default_rng(7).choice(np.nonzero(g.out_degree > 0)[0], 64,
                                                                      replace=False)]
            for _ in range(2):
                G.msbfs(arena, g, srcs64, sync=False)
                G.mssp(arena, gw, srcs64, sync=False)
    elif what == "barrier":
        # the headline launches: 1024 WGs x 256, 16 stores per thread, 1e5
        # episodes, the XCD barrier with the fan-in hs_tune_barrier picks (as the
        # bench applies it), and the cooperative grid sync
        storage = torch.zeros(1024 * 256 * 17, dtype=torch.float32, device="cuda")
        groups = hs.tune_barrier(arena, storage, 1024, 256, 16, episodes=10000, reps=2)["xcd_groups"]
        hs.set_xcd_groups(groups)
        hs.barrier_bench(arena, "xcd", storage, 1024, 256, 16, 100000, check=False, sync=False)
        arena.sync()
        hs.set_xcd_groups(0)
        hs.barrier_bench(arena, "coop", storage, 1024, 256, 16, 100000, check=False, sync=False)
    else:
        # counter runs: 2000 episodes per launch, with and without the data
        # pattern, so that per-episode counts of the synchronization alone and
        # of the whole episode can be told apart
        storage = torch.zeros(2048 * 256 * 17, dtype=torch.float32, device="cuda")
        for kind, ldst, wgs in BARRIER_COUNTER_RUNS:
            hs.barrier_bench(arena, kind, storage, wgs, 256, ldst, 2000, check=False, sync=False)
    arena.sync()
    torch.cuda.synchronize()
    arena.close()
    print(f"profile {what}: done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--episodes", type=int, default=100000)
    ap.add_argument("--no-extras", action="store_true", help="headline barrier only")
    ap.add_argument("--bfs-scale", type=int, default=22)
    ap.add_argument("--profile", choices=["reduce", "scan", "barrier", "barrier_counters", "pagerank", "graph"],
                    default=None,
                    help="only launch one kernel family (for rocprofv3 runs), print nothing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the synchronization "
              "primitives are per-device and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch
    import hsync as hs
    hs.load()
    if args.profile:
        profile_only(torch, hs, args.profile)
        return
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

    # the barrier's fan-in (XCD groups) picked by the library's autotuner on this
    # workload, applied for the headline; every other row keeps the default
    tune = hs.tune_barrier(arena, storage, W, WG, L, episodes=10000, reps=2)
    groups = tune["xcd_groups"]
    log("headline: xcd barrier", W, "WGs, tuned groups", groups, tune)
    hs.set_xcd_groups(groups)
    ms_step, ns_xcd = measure_barrier("xcd")
    hs.set_xcd_groups(0)
    log(f"  xcd: {ns_xcd:.1f} ns/episode ({ms_step:.2f} ms/step)")
    extra["barrier_tune"] = {"xcd_groups": groups, "tuned_ns_per_episode_10k": round(tune["ms"] * 1e6 / 10000, 1),
                             "default_groups_ns_per_episode_10k": round(tune["default_ms"] * 1e6 / 10000, 1),
                             "tried": tune["tried"]}
    results = {"xcd": ns_xcd}
    if not args.no_extras:
        for kind in ("coop", "two_level", "flat", "xcd_flat_release", "lf", "tree"):
            eps = E // 5 if kind in ("coop", "two_level", "flat", "xcd_flat_release") else E // 10
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
                    what = {3: "refused (grid not co-resident)", 5: "timeout (a wait exceeded the spin limit)"}
                    row[kind] = f"{what.get(ex.status, 'error')}: {ex}"
            sweep[str(wgs)] = row
            log("  sweep", wgs, row)
        extra["barrier_sweep_ns_per_episode"] = sweep
        # semaphores: 256 / 512 / 1024 / 2048 WGs x 256, 16 loads + 16 stores
        # in the critical section, readers:writers 9:1, capacity 1 / 16 / 256.
        # A configuration where some workgroup waits longer than the arena's
        # spin limit (starvation of the unfair kinds) is reported as such.  The
        # unfair kinds' cost per acquisition grows with the contenders (spin:
        # ~0.8 ms at 256 WGs), so they run fewer iterations at larger grids and
        # not at 2048.
        sem_names = {0: "spin", 1: "ebo", 2: "spin_priority", 3: "ebo_priority", 4: "xcd_queue"}
        sem_plan = ((256, 20, 5, (0, 1, 2, 3, 4)), (512, 10, 2, (0, 1, 2, 3, 4)), (1024, 5, 1, (0, 1, 2, 3, 4)),
                    (2048, 2, 1, (2, 3, 4)))
        sem_res = {}
        for wgs, iters, iters_unfair, kinds in sem_plan:
            for kind in kinds:
                name = sem_names[kind]
                it = iters_unfair if kind in (0, 1) else iters
                row = {}
                for cap in (1, 16, 256):
                    t_start = time.perf_counter()
                    try:
                        sem = arena.sem_create(kind, kind, cap)
                        v = torch.zeros(128, dtype=torch.int32, device="cuda")
                        if kind in (2, 3, 4):  # warm-up of the fast kinds
                            hs.sem_bench(arena, kind, sem, storage, wgs, WG, L, 1, 10)
                        ms = timed_launches(torch, lambda i: hs.sem_bench(arena, kind, sem, storage, wgs, WG, L,
                                                                          it, 10, violations=v, sync=False), 1)
                        arena.sync()
                        row[str(cap)] = {"ns_per_acquire": round(ms * 1e6 / (wgs * it), 1),
                                         **hs.sem_wait_stats(v, wgs * it)}
                    except hs.HSError as ex:
                        row[str(cap)] = f"wait exceeded the spin limit ({ex.status})"
                    log(f"  semaphore {name} {wgs} WGs capacity {cap}: {row[str(cap)]} "
                        f"({time.perf_counter() - t_start:.1f} s)")
                sem_res[f"{name}@{wgs}"] = row
        extra["semaphore_9to1"] = sem_res
        # bounded wait of the per-XCD ticket queue: 1024 WGs x 977 acquisitions
        # (1,000,448 per capacity), 9:1 mix, every acquisition checked and timed
        qrun = {}
        for cap in (1, 16, 256):
            sem = arena.sem_create(4, 4, cap)
            v = torch.zeros(128, dtype=torch.int32, device="cuda")
            t_start = time.perf_counter()
            hs.sem_bench(arena, 4, sem, storage, 1024, WG, L, 977, 10, violations=v)
            sec = time.perf_counter() - t_start
            qrun[str(cap)] = {"seconds": round(sec, 2), "ns_per_acquire": round(sec * 1e9 / (1024 * 977), 1),
                              **hs.sem_wait_stats(v, 1024 * 977)}
            log("  xcd_queue 1e6 acquisitions, capacity", cap, qrun[str(cap)])
        extra["xcd_queue_1e6_acquisitions_1024wg"] = qrun
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

    # K launches captured in one hipGraph, so the timed region holds back-to-back
    # kernels (no Python launch overhead between them)
    from hsync.graphcap import Replay
    K = 64
    g_cold = Replay(lambda: [red(i) for i in range(K)])
    g_warm = Replay(lambda: [red(i, True) for i in range(K)])
    warm_clocks(torch, lambda i: g_cold.replay())
    arena.sync()
    R = 20
    ms_cold = timed_launches(torch, lambda i: g_cold.replay(), R) / K
    ms_warm = timed_launches(torch, lambda i: g_warm.replay(), R) / K
    arena.sync()
    del g_cold, g_warm
    # the configuration BASELINE names (2 WGs per CU), and the launch + fan-in
    # alone (a reduce of 0 elements)
    g2 = Replay(lambda: [hs.reduce_i32(arena, bufs[i % nbuf], barrier="xcd", out=out, partials=partials,
                                       num_wgs=2 * cus, sync=False) for i in range(K)])
    g0 = Replay(lambda: [hs.reduce_i32(arena, bufs[0][:0], barrier="xcd", out=out, partials=partials,
                                       sync=False) for i in range(K)])
    warm_clocks(torch, lambda i: g2.replay())
    ms_cold2 = timed_launches(torch, lambda i: g2.replay(), R) / K
    ms_empty = timed_launches(torch, lambda i: g0.replay(), R) / K
    arena.sync()
    del g2, g0
    # in-kernel clock stamps of the same HBM-cold launches (100 MHz): dispatch
    # spread, streaming and fold inside a launch, against the launch-to-launch time
    import numpy as np
    stamps = torch.zeros(K * 4 * cus, dtype=torch.int64, device="cuda")
    gs = Replay(lambda: [hs.reduce_i32_stamped(arena, bufs[i % nbuf], stamps[i * 4 * cus:(i + 1) * 4 * cus],
                                               barrier="xcd", out=out, partials=partials) for i in range(K)])
    warm_clocks(torch, lambda i: gs.replay())
    gs.replay()
    torch.cuda.synchronize()
    arena.sync()
    parts = [hs.stamp_breakdown(stamps[i * 4 * cus:(i + 1) * 4 * cus], cus) for i in range(K)]
    in_kernel = {k: round(float(np.median([p_[k] for p_ in parts])), 3) for k in parts[0]}
    del gs
    # correctness of the timed kernel on buffer 0
    s_gpu = hs.reduce_i32(arena, bufs[0])
    s_ref = int(bufs[0].sum(dtype=torch.int64).item())
    assert (s_gpu - s_ref) % 2**32 == 0, (s_gpu, s_ref)  # int32 wrap
    gbps_cold = 4 * N / (ms_cold * 1e-3) / 1e9
    gbps_warm = 4 * N / (ms_warm * 1e-3) / 1e9
    log(f"reduce 2^24: cold {ms_cold*1e3:.2f} us {gbps_cold:.0f} GB/s, warm {ms_warm*1e3:.2f} us "
        f"{gbps_warm:.0f} GB/s")
    extra["reduce_2p24"] = {"GBps_hbm_cold": round(gbps_cold, 1), "GBps_mall_warm": round(gbps_warm, 1),
                            "us_per_launch_cold": round(ms_cold * 1e3, 2),
                            "us_per_launch_warm": round(ms_warm * 1e3, 2),
                            "grid": f"{cus} WGs x 256 (1 per CU)",
                            "GBps_hbm_cold_2_per_cu": round(4 * N / (ms_cold2 * 1e-3) / 1e9, 1),
                            "us_per_launch_empty": round(ms_empty * 1e3, 2),
                            "in_kernel_median_us": in_kernel,
                            "GBps_in_kernel_stream": round(4 * N / (in_kernel["stream_us"] * 1e-6) / 1e9, 1),
                            "barrier": "xcd (arrival half: the last arriver folds)", "sum_matches": True}
    log("  reduce in-kernel stamps", in_kernel)
    del bufs

    cpu_scan, cpu_graph = {}, {}
    # ---------------- scan: 2^28 int32 uniform [0,15], inclusive
    if not args.no_extras:
        NS = 1 << 28
        x = torch.empty(NS, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        hs.gen_uniform_u32(x, 0, 0, 15)
        scratch = torch.empty(8192, dtype=torch.int32, device="cuda")
        warm_clocks(torch, lambda i: hs.scan_u32(arena, x, inclusive=True, out=y, scratch=scratch,
                                                  sync=False))
        arena.sync()
        ms_scan = timed_launches(torch, lambda i: hs.scan_u32(arena, x, inclusive=True, out=y,
                                                              scratch=scratch, sync=False), 5)
        arena.sync()
        total = int(x.sum(dtype=torch.int64).item()) & 0xFFFFFFFF
        ok = (int(y[-1].item()) & 0xFFFFFFFF) == total
        gb_scan = 8 * NS / (ms_scan * 1e-3) / 1e9
        # practical ceilings of this box for the same bytes: a 1 GiB device copy
        # (read + write, what a scan must do) and a 1 GiB read (our reduce)
        for _ in range(5):  # warm: the first copies after other work read slow
            y.copy_(x)
        ms_copy = timed_launches(torch, lambda i: y.copy_(x), 10)
        ms_read = timed_launches(torch, lambda i: hs.reduce_i32(arena, x, sync=False), 5)
        arena.sync()
        copy_gbps = 8 * NS / (ms_copy * 1e-3) / 1e9
        read_gbps = 4 * NS / (ms_read * 1e-3) / 1e9
        extra["hbm_practical_ceilings"] = {"copy_1GiB_GBps": round(copy_gbps, 1),
                                           "read_1GiB_GBps": round(read_gbps, 1)}
        extra["scan_2p28_inclusive"] = {"method": "multi-phase barrier scan (XCD barrier, pipelined rounds)",
                                        "GBps": round(gb_scan, 1), "ms": round(ms_scan, 3),
                                        "roofline_frac": round(gb_scan / HBM_PEAK_GBPS, 3),
                                        "frac_of_copy_ceiling": round(gb_scan / copy_gbps, 3),
                                        "last_equals_total": ok}
        log("scan 2^28:", extra["scan_2p28_inclusive"], extra["hbm_practical_ceilings"])
        # the same scan as one segment of the single-pass look-back kernel
        import numpy as np
        o1 = np.array([0, NS], dtype=np.uint64)
        import ctypes as C
        nb = C.c_uint64(0)
        hs.load().hs_batch_work_size(o1.ctypes.data, 1, C.byref(nb))
        w1 = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
        hs.scan_batch_u32(x, o1, out=y, work=w1)
        warm_clocks(torch, lambda i: hs.scan_batch_u32(x, o1, out=y, work=w1, sync=False))
        ms_lb = timed_launches(torch, lambda i: hs.scan_batch_u32(x, o1, out=y, work=w1, sync=False), 10)
        hs.scan_batch_u32(x, o1, out=y, work=w1)
        gb_lb = 8 * NS / (ms_lb * 1e-3) / 1e9
        extra["scan_2p28_lookback"] = {"method": "single-pass decoupled look-back (wave-parallel)",
                                       "GBps": round(gb_lb, 1), "ms": round(ms_lb, 3),
                                       "roofline_frac": round(gb_lb / HBM_PEAK_GBPS, 3),
                                       "frac_of_copy_ceiling": round(gb_lb / copy_gbps, 3),
                                       "last_equals_total": (int(y[-1].item()) & 0xFFFFFFFF) == total}
        log("scan 2^28 look-back:", extra["scan_2p28_lookback"])
        cpu_scan.update(cpu_scan_leg(x.cpu().numpy(), y.cpu().numpy()))
        log("  cpu scan", cpu_scan)
        del x, y

    if not args.no_extras:
        extra["batched_scan_512"] = bench_batched_scan(torch, hs)
        log("batched scan:", extra["batched_scan_512"])
        extra["bfs_rmat22_64src"] = bench_bfs(torch, hs, arena, args.bfs_scale, cpu_graph)
        log("bfs:", extra["bfs_rmat22_64src"])
        extra["serving"] = bench_serving(torch, hs)
        log("serving:", extra["serving"])

    # ---------------- CPU baseline (host cores, oracle library)
    cpu = cpu_baseline_leg(N)
    if cpu_scan:
        cpu["scan_2p28"] = cpu_scan
    if cpu_graph:
        cpu["graph_rmat22"] = cpu_graph
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
                               "XCD-hierarchical barrier (fan-in from hs_tune_barrier)",
                   "wgs": W, "wg_size": WG, "stores_per_thread": L, "episodes": E,
                   "xcd_groups": groups},
        "roofline": {"kernel": "k_reduce_i32 (2^24 int32, HBM-cold rotation)", "bound": "hbm",
                     "achieved": round(gbps_cold, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps_cold / HBM_PEAK_GBPS, 4), "traffic": profiled_traffic(),
                     "algorithmic_bytes": 4 * N,
                     "frac_of_read_ceiling": (round(gbps_cold / extra["hbm_practical_ceilings"]["read_1GiB_GBps"], 4)
                                              if "hbm_practical_ceilings" in extra else None)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    arena.close()


if __name__ == "__main__":
    main()
