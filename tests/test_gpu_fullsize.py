"""Parity at the full sizes of the headline configurations, against the CPU
oracle: 64-source BFS and SSSP on R-MAT scale 22 (4,194,304 vertices, ~65M
edges), PageRank on the same graph against the FP64 oracle, and the serving
loop on a 2^26-element reduce and scan and on R-MAT 20 BFS and PageRank to an
L1 change below 1e-6."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat22():
    import torch  # noqa: F401  (one HIP runtime: torch first)
    import hsync.graph as G
    _log("building R-MAT 22")
    g = G.Graph.rmat(22, 16, 1)
    deg = g.out_degree
    srcs = [int(v) for v in np.random.default_rng(7).choice(np.nonzero(deg > 0)[0], 64, replace=False)]
    return G, g, srcs


def test_msbfs_rmat22_all_64_rows(hs, arena, oracle, rmat22):
    G, g, srcs = rmat22
    ref, threads = oracle.cpu_msbfs(g.h_row, g.h_col, g.h_rrow, g.h_rcol, srcs)
    for sched in ("heuristic", "push", "pull"):
        _log("msbfs R-MAT 22", sched)
        lv = G.msbfs(arena, g, srcs, schedule=sched).cpu().numpy()
        bad = [i for i in range(64) if not np.array_equal(lv[i], ref[i])]
        assert not bad, (sched, bad[:8])


def test_mssp_rmat22_sampled_rows(hs, arena, oracle, rmat22):
    G, g, srcs = rmat22
    w = G.weights(g.m, 1001)  # uniform uint32 [1, 255], the bench's weights
    gw = G.Graph(g.h_row, g.h_col, w, transposed=False)
    sample = [0, 9, 18, 27, 36, 45, 54, 63]
    ref, _ = oracle.cpu_mssp(g.h_row, g.h_col, w, [srcs[i] for i in sample])
    for delta in (0x7FFFFFFF, 10000):
        _log("mssp R-MAT 22 delta", delta)
        dist, _ = G.mssp(arena, gw, srcs, delta=delta)
        d = dist.cpu().numpy()
        for j, i in enumerate(sample):
            assert np.array_equal(d[i], ref[j]), (delta, i)
    # the plain persistent near/far SSSP (one source per launch) on two rows
    for j in (0, 7):
        d1, _ = G.sssp(arena, gw, srcs[sample[j]])
        assert np.array_equal(d1.cpu().numpy(), ref[j]), sample[j]


def test_pagerank_rmat22_vs_fp64(hs, arena, oracle, rmat22):
    G, g, _ = rmat22
    rank, st = G.pagerank(arena, g, tol=1e-6)
    it = int(st[0].item())
    got = rank.cpu().numpy().astype(np.float64)
    # the FP64 oracle: the same normalized iteration converged much further
    ref, it_ref = oracle.pagerank(g.h_row, g.h_col, alpha=0.85, tol=1e-12)
    err = float(np.abs(got - ref).max())
    _log(f"pagerank R-MAT 22: {it} FP32 iterations, max-abs error vs FP64 {err:.3e} "
         f"(FP64 oracle {it_ref} iterations), rank sum {got.sum():.6f}")
    assert err <= 1e-6
    # identical bits for every barrier kind (deterministic fixed-point partials)
    for kind in ("two_level", "flat"):
        r2, _ = G.pagerank(arena, g, tol=1e-6, barrier=kind)
        assert hs.api._torch().equal(r2, rank), kind


def test_serve_full_size_requests(hs, oracle):
    """One 2^26 reduce, one 2^26 scan, and BFS plus PageRank to an L1 change
    below 1e-6 on R-MAT 20, each alone and then together, against the oracle."""
    import torch
    import hsync.graph as G
    from hsync.serve import Server
    n = 1 << 26
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 260, 0, 255)
    hx = oracle.gen_uniform_u32(260, n, 0, 255)
    row, col = G.rmat_csr(20, 16, 20)
    g = G.Graph(row, col)
    deg = np.diff(row.astype(np.int64))
    src = int(np.random.default_rng(20).choice(np.nonzero(deg > 0)[0]))
    outs = {"reduce": torch.zeros(1, dtype=torch.int32, device="cuda"),
            "scan": torch.zeros(n, dtype=torch.int32, device="cuda"),
            "bfs": torch.zeros(g.n, dtype=torch.int32, device="cuda"),
            "pr": torch.zeros(g.n, dtype=torch.float32, device="cuda")}
    bfs_s = torch.zeros(2 * g.n, dtype=torch.int32, device="cuda")
    pr_s = torch.zeros(2 * g.n, dtype=torch.float32, device="cuda")
    exp_red = oracle.reduce_i32(hx.astype(np.int32))
    exp_scan = oracle.scan_u32(hx, True)
    exp_bfs = oracle.bfs(row, col, src)
    exp_pr, it_ref = oracle.pagerank(row, col, alpha=0.85, tol=1e-6, max_iter=1000)
    torch.cuda.synchronize()
    srv = Server(num_wgs=512, ring=16, budget_units=64)
    try:
        one = {"reduce": lambda: srv.reduce(x, outs["reduce"]), "scan": lambda: srv.scan(x, outs["scan"]),
               "bfs": lambda: srv.bfs(g, src, outs["bfs"], bfs_s),
               "pr": lambda: srv.pagerank(g, 1000, outs["pr"], pr_s, tol=1e-6)}

        def check(tag, idx):
            for k, i in idx.items():
                assert srv.wait(i, timeout=120)["status"] == 0, (tag, k)
            got = {k: outs[k].cpu().numpy().copy() for k in idx}
            if "reduce" in got:
                assert int(got["reduce"][0]) == exp_red, tag
            if "scan" in got:
                assert np.array_equal(got["scan"].view(np.uint32), exp_scan), tag
            if "bfs" in got:
                assert np.array_equal(got["bfs"], exp_bfs), tag
            if "pr" in got:
                err = float(np.abs(got["pr"].astype(np.float64) - exp_pr).max())
                _log(f"serve {tag}: pagerank R-MAT 20 {srv.phases(idx['pr']) - 1} iterations "
                     f"(oracle {it_ref}), max-abs error {err:.3e}")
                assert err < 1e-6, tag
            return got

        alone = {}
        for k, f in one.items():
            alone.update(check("alone " + k, {k: f()}))
        mix = check("mixed", {k: f() for k, f in one.items()})
        for k in alone:
            assert np.array_equal(alone[k], mix[k]), k
    finally:
        srv.stop()
