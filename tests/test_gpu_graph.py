"""GPU parity of the persistent graph kernels against the CPU oracle."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
INF = 2**31 - 1
KINDS = ["flat", "two_level", "xcd", "xcd_flat_release", "tree", "lf", "coop", "naive_all"]


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hg():
    import torch  # noqa: F401  (one HIP runtime: torch first)
    import hsync.graph as g
    return g


def _small_graphs(hg):
    gs = {}
    # path 0->1->...->99
    row = np.array(list(range(100)) + [99], np.uint32)
    gs["path"] = (row, np.arange(1, 100, dtype=np.uint32))
    # star + isolated vertices
    row = np.array([0, 50] + [50] * 60, np.uint32)
    gs["star"] = (row, np.arange(1, 51, dtype=np.uint32))
    r, c, _ = hg.read_gr(os.path.join(HERE, "golden", "tiny.gr"))
    gs["tiny"] = (r, c)
    gs["rmat10"] = hg.rmat_csr(10, 16, 1)
    gs["rmat12"] = hg.rmat_csr(12, 16, 7)
    return gs


def test_bfs_small_and_rmat(hs, arena, oracle, hg):
    for name, (row, col) in _small_graphs(hg).items():
        g = hg.Graph(row, col)
        for src in sorted({0, 1, g.n // 2, g.n - 1}):
            _log("bfs", name, src)
            lv, st = hg.bfs(arena, g, src)
            assert np.array_equal(lv.cpu().numpy(), oracle.bfs(row, col, src)), (name, src)


@pytest.mark.parametrize("kind", KINDS)
def test_bfs_all_barriers(hs, arena, oracle, hg, kind):
    row, col = hg.rmat_csr(14, 16, 3)
    g = hg.Graph(row, col)
    exp = oracle.bfs(row, col, 0)
    for wgs in (0, 64):
        lv, _ = hg.bfs(arena, g, 0, barrier=kind, num_wgs=wgs)
        assert np.array_equal(lv.cpu().numpy(), exp), (kind, wgs)


def _sources(row, k, seed):
    deg = np.diff(row.astype(np.int64))
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(cand, size=min(k, cand.size), replace=False)]


@pytest.mark.parametrize("schedule", ["heuristic", "push", "pull", "forced"])
def test_msbfs_matches_single_source(hs, arena, oracle, hg, schedule):
    import torch
    row, col = hg.rmat_csr(12, 16, 7)
    g = hg.Graph(row, col)
    srcs = _sources(row, 64, 5)
    forced = list(np.random.default_rng(3).integers(0, 2, 64)) if schedule == "forced" else None
    tr = hg.new_trace()
    lv = hg.msbfs(arena, g, srcs, schedule=schedule, forced=forced, trace=tr).cpu().numpy()
    for i, s in enumerate(srcs):
        assert np.array_equal(lv[i], oracle.bfs(row, col, s)), (schedule, i, s)
    recs = hg.decode_trace(tr)
    assert len(recs) >= 2
    if schedule == "push":
        assert all(r["dir"] == 0 for r in recs)
    if schedule == "pull":
        assert all(r["dir"] == 1 for r in recs)
    # fewer sources, ragged masks, duplicate sources
    lv2 = hg.msbfs(arena, g, srcs[:5] + [srcs[0]], schedule=schedule, forced=forced).cpu().numpy()
    assert np.array_equal(lv2[:5], lv[:5]) and np.array_equal(lv2[5], lv[0])
    del torch


@pytest.mark.parametrize("kind", ["flat", "two_level", "xcd", "lf", "coop"])
def test_msbfs_barriers(hs, arena, oracle, hg, kind):
    row, col = hg.rmat_csr(13, 16, 11)
    g = hg.Graph(row, col)
    srcs = _sources(row, 64, 2)
    a = hg.msbfs(arena, g, srcs, barrier=kind).cpu().numpy()
    for i in (0, 17, 63):
        assert np.array_equal(a[i], oracle.bfs(row, col, srcs[i]))


def test_sssp(hs, arena, oracle, hg):
    for scale, seed in ((10, 1), (12, 4)):
        row, col = hg.rmat_csr(scale, 16, seed)
        w = hg.weights(col.size, seed + 100)
        g = hg.Graph(row, col, w)
        for src in _sources(row, 3, seed):
            exp = oracle.sssp(row, col, w, src)
            for delta in (10000, 64, 1):
                _log("sssp", scale, src, delta)
                d, st = hg.sssp(arena, g, src, delta=delta)
                assert np.array_equal(d.cpu().numpy(), exp), (scale, src, delta)
    # unit weights: SSSP == BFS levels
    row, col = hg.rmat_csr(11, 8, 2)
    g = hg.Graph(row, col, np.ones(col.size, np.uint32))
    d, _ = hg.sssp(arena, g, 0)
    assert np.array_equal(d.cpu().numpy(), oracle.bfs(row, col, 0))


@pytest.mark.parametrize("kind", ["flat", "xcd", "lf", "coop"])
def test_sssp_barriers(hs, arena, oracle, hg, kind):
    row, col = hg.rmat_csr(12, 16, 9)
    w = hg.weights(col.size, 5)
    g = hg.Graph(row, col, w)
    d, _ = hg.sssp(arena, g, 3, delta=100, barrier=kind)
    assert np.array_equal(d.cpu().numpy(), oracle.sssp(row, col, w, 3))


@pytest.mark.parametrize("delta", [0x7FFFFFFF, 1000, 64])
def test_mssp_matches_dijkstra_and_single_source(hs, arena, oracle, hg, delta):
    row, col = hg.rmat_csr(12, 16, 21)
    w = hg.weights(col.size, 22)
    g = hg.Graph(row, col, w)
    srcs = _sources(row, 64, 23)
    srcs[5] = srcs[4]  # the same vertex twice
    _log("mssp", delta)
    d, st = hg.mssp(arena, g, srcs, delta=delta)
    d = d.cpu().numpy()
    for i in (0, 4, 5, 17, 40, 63):
        assert np.array_equal(d[i], oracle.sssp(row, col, w, srcs[i])), (delta, i)
    for i in (1, 33):
        single, _ = hg.sssp(arena, g, srcs[i], delta=100)
        assert np.array_equal(d[i], single.cpu().numpy()), (delta, i)
    assert int(st[0].item()) > 0


@pytest.mark.parametrize("kind", ["flat", "two_level", "xcd", "lf", "coop"])
def test_mssp_barriers_and_small_batches(hs, arena, oracle, hg, kind):
    row, col = hg.rmat_csr(11, 8, 31)
    w = hg.weights(col.size, 32)
    g = hg.Graph(row, col, w)
    for nsrc in (1, 7):
        srcs = _sources(row, nsrc, 33 + nsrc)
        d, _ = hg.mssp(arena, g, srcs, delta=256, barrier=kind)
        d = d.cpu().numpy()
        assert d.shape == (nsrc, g.n)
        for i in range(nsrc):
            assert np.array_equal(d[i], oracle.sssp(row, col, w, srcs[i])), (kind, nsrc, i)
    # unit weights: every row equals the BFS levels
    g1 = hg.Graph(row, col, np.ones(col.size, np.uint32))
    srcs = _sources(row, 5, 35)
    d, _ = hg.mssp(arena, g1, srcs, barrier=kind)
    d = d.cpu().numpy()
    for i in range(5):
        assert np.array_equal(d[i], oracle.bfs(row, col, srcs[i]))


def test_pagerank_vs_fp64(hs, arena, oracle, hg):
    for scale, seed in ((10, 1), (13, 2)):
        row, col = hg.rmat_csr(scale, 16, seed)
        g = hg.Graph(row, col)
        ref, _ = oracle.pagerank(row, col, alpha=0.85, tol=1e-14)
        r, st = hg.pagerank(arena, g, alpha=0.85, tol=1e-6)
        x = r.cpu().numpy().astype(np.float64)
        err = np.abs(x - ref).max()
        _log("pagerank", scale, "iters", int(st[0].item()), "max abs err", err)
        assert err <= 1e-6
        assert abs(x.sum() - 1.0) < 1e-4
        # deterministic: same bits on a re-run and with another barrier kind
        r2, _ = hg.pagerank(arena, g, alpha=0.85, tol=1e-6, barrier="flat")
        assert np.array_equal(r.cpu().numpy().view(np.uint32), r2.cpu().numpy().view(np.uint32))


def test_pagerank_closed_form_cycle(hs, arena, hg):
    n = 1000
    row = np.arange(n + 1, dtype=np.uint32)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)
    r, _ = hg.pagerank(arena, hg.Graph(row, col))
    assert np.abs(r.cpu().numpy() - 1.0 / n).max() < 1e-9
