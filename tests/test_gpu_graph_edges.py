"""Edge cases of the persistent graph kernels against the CPU oracle: a single
vertex, isolated vertices, a source without out-edges, self loops and
duplicate edges, hubs past every heavy-vertex threshold (out-degree and
in-degree 5000: grid-shared BFS/SSSP chunks, PageRank chunk slots), and a path
of 600 levels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
INF = 2**31 - 1


def _csr(n, edges):
    edges = sorted(edges)
    row = np.zeros(n + 1, np.uint32)
    for u, _ in edges:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.uint32)
    col = np.array([v for _, v in edges], np.uint32)
    return row, col


def _cases():
    rng = np.random.default_rng(9)
    cases = {"single": _csr(1, []), "isolated": _csr(5, [(1, 2)])}
    e = [(0, 0), (0, 1), (0, 1), (1, 2), (2, 2), (2, 0), (1, 2)]
    cases["loops_dups"] = _csr(3, e)
    hub = [(0, v) for v in range(1, 5001)] + [(v, 5001) for v in range(1, 5001)]
    hub += [(int(a), int(b)) for a, b in rng.integers(1, 5002, (3000, 2))]
    cases["hubs"] = _csr(5002, hub)
    # a path deeper than 255 levels (the batched BFS writes such levels directly)
    cases["path600"] = _csr(600, [(v, v + 1) for v in range(599)])
    return cases


@pytest.fixture(scope="module")
def hg():
    import torch  # noqa: F401  (one HIP runtime: torch first)
    import hsync.graph as g
    return g


def test_bfs_sssp_edge_cases(hs, arena, oracle, hg):
    for name, (row, col) in _cases().items():
        n = row.size - 1
        w = (np.arange(col.size, dtype=np.uint32) % 7 + 1).astype(np.uint32)
        g = hg.Graph(row, col, w if col.size else None)
        for src in sorted({0, n - 1, n // 2}):
            print("edge case", name, "src", src, flush=True)
            lv, _ = hg.bfs(arena, g, src)
            assert np.array_equal(lv.cpu().numpy(), oracle.bfs(row, col, src)), (name, src)
            if col.size:
                d, _ = hg.sssp(arena, g, src, delta=5)
                assert np.array_equal(d.cpu().numpy(), oracle.sssp(row, col, w, src)), (name, src)


def test_batched_edge_cases(hs, arena, oracle, hg):
    for name, (row, col) in _cases().items():
        n = row.size - 1
        if not col.size:
            continue
        w = (np.arange(col.size, dtype=np.uint32) % 7 + 1).astype(np.uint32)
        g = hg.Graph(row, col, w)
        srcs = sorted({0, n - 1, n // 2, 1 % n})
        print("batched edge case", name, srcs, flush=True)
        lv = hg.msbfs(arena, g, srcs).cpu().numpy()
        d, _ = hg.mssp(arena, g, srcs, delta=5)
        d = d.cpu().numpy()
        for i, s in enumerate(srcs):
            assert np.array_equal(lv[i], oracle.bfs(row, col, s)), (name, s)
            assert np.array_equal(d[i], oracle.sssp(row, col, w, s)), (name, s)


def test_pagerank_edge_cases(hs, arena, oracle, hg):
    for name, (row, col) in _cases().items():
        print("pagerank edge case", name, flush=True)
        g = hg.Graph(row, col)
        ref, _ = oracle.pagerank(row, col, alpha=0.85, tol=1e-14)
        r, st = hg.pagerank(arena, g, alpha=0.85, tol=1e-7)
        x = r.cpu().numpy().astype(np.float64)
        assert np.abs(x - ref).max() < 1e-6, name
        assert abs(x.sum() - 1.0) < 1e-4, name


def test_bfs_load_balanced(hs, arena, oracle, hg):
    """hs_bfs_lb (the reference's -l): frontier vertices of out-degree >= the
    limit have their edges split evenly over the grid.  Levels equal the
    oracle's for limits that send hubs (degree 5000), R-MAT hubs or nothing to
    the balanced path (the default limit, the grid's thread count)."""
    cases = dict(_cases())
    cases["rmat14"] = hg.rmat_csr(14, 16, 77)
    for name, (row, col) in cases.items():
        n = row.size - 1
        g = hg.Graph(row, col, transposed=False)
        srcs = sorted({0, n - 1, n // 2, int(np.argmax(np.diff(row.astype(np.int64))))})
        for lb in (256, 1000, 0):
            for barrier in ("xcd", "coop"):
                for src in srcs:
                    print("lb", name, lb, barrier, src, flush=True)
                    lv, st = hg.bfs(arena, g, src, barrier=barrier, lb=lb)
                    assert np.array_equal(lv.cpu().numpy(), oracle.bfs(row, col, src)), (name, lb, barrier, src)


def test_bfs_load_balanced_rejects_small_limit(hs, arena, hg):
    row, col = _cases()["hubs"]
    g = hg.Graph(row, col, transposed=False)
    with pytest.raises(hs.HSError):
        hg.bfs(arena, g, 0, lb=255)


def test_msbfs_sources_in_disjoint_components(hs, arena, oracle, hg):
    """Sources spread over three components of different depths: a source's
    search ends while the others go on, and the pull must treat its bit as
    found (the vertices of the other components never get it)."""
    r1, c1 = hg.rmat_csr(11, 8, 5)
    n1 = r1.size - 1
    path = [(v, v + 1) for v in range(40)] + [(v + 1, v) for v in range(40)]
    edges = [(int(u), int(w)) for u in range(n1) for w in c1[r1[u]:r1[u + 1]]]
    edges += [(n1 + a, n1 + b) for a, b in path]  # a path of 41 vertices
    edges += [(n1 + 41, n1 + 42), (n1 + 42, n1 + 41)]  # a pair
    row, col = _csr(n1 + 43, edges)
    g = hg.Graph(row, col)
    deg = np.diff(r1.astype(np.int64))
    big = [int(v) for v in np.random.default_rng(3).choice(np.nonzero(deg > 0)[0], 60, replace=False)]
    srcs = big[:30] + [n1, n1 + 20] + big[30:] + [n1 + 41, n1 + 40]
    for sched in ("heuristic", "pull", "push"):
        lv = hg.msbfs(arena, g, srcs, schedule=sched).cpu().numpy()
        for i in range(len(srcs)):
            assert np.array_equal(lv[i], oracle.bfs(row, col, srcs[i])), (sched, i, srcs[i])
