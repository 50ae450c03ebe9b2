"""Pins the CPU oracle against published vectors, closed forms, invariants and
independent library routines (scipy) -- not against a retyping of itself."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- generator
def test_splitmix64_published_vector(oracle):
    fx = json.load(open(os.path.join(HERE, "golden", "splitmix64.json")))
    # Vigna's splitmix64.c, state 1234567: output k = mix64(1234567 + k*golden)
    G = 0x9E3779B97F4A7C15
    got = [oracle.mix64((1234567 + k * G) & ((1 << 64) - 1)) for k in range(1, 6)]
    assert got == fx["splitmix64_state_1234567"]
    assert got[0] == 6457827717110365317


def test_uniform_generator_fixture_and_range(oracle):
    fx = json.load(open(os.path.join(HERE, "golden", "splitmix64.json")))["uniform_u32"]
    for key, vals in fx.items():
        seed, lo, hi = (int(x) for x in key.split(":"))
        assert oracle.gen_uniform_u32(seed, 16, lo, hi).tolist() == vals
    a = oracle.gen_uniform_u32(7, 1 << 20, 0, 255)
    assert a.min() == 0 and a.max() == 255
    # roughly uniform: every value appears ~4096 times
    h = np.bincount(a, minlength=256)
    assert h.min() > 3500 and h.max() < 4700


# ---------------------------------------------------------------- reduce / scan
def test_reduce_closed_forms(oracle):
    n = 81920  # Reduce/README.txt:14 max-contention size
    assert oracle.reduce_i32(np.ones(n, np.int32)) == n  # Reduce/test_reduce.cu:14-17
    r = np.arange(100000, dtype=np.int32)
    assert oracle.reduce_i64(r) == 100000 * 99999 // 2
    a = np.random.default_rng(1).integers(-2**31, 2**31, 50000, dtype=np.int64).astype(np.int32)
    assert oracle.reduce_i64(a) == int(a.astype(np.int64).sum())
    # int32 wrap: the reference sums with int arithmetic
    wrapped = ((int(a.astype(np.int64).sum()) + 2**31) % 2**32) - 2**31
    assert oracle.reduce_i32(a) == wrapped
    assert oracle.reduce_i32(np.zeros(0, np.int32)) == 0


def test_scan_closed_forms(oracle):
    n = 163840  # scan/README.txt:22
    ones = np.ones(n, np.uint32)  # scan/main.cu:36-41 input
    assert np.array_equal(oracle.scan_u32(ones, False), np.arange(n, dtype=np.uint32))
    assert np.array_equal(oracle.scan_u32(ones, True), np.arange(1, n + 1, dtype=np.uint32))
    a = np.random.default_rng(2).integers(0, 2**32, 100001, dtype=np.uint64).astype(np.uint32)
    inc = oracle.scan_u32(a, True)
    exc = oracle.scan_u32(a, False)
    assert np.array_equal(inc, np.cumsum(a, dtype=np.uint32))
    assert np.array_equal(inc, exc + a)  # uint32 wrap on both sides


def test_cpu_baselines_agree(oracle):
    a = oracle.gen_uniform_u32(3, 1 << 20, 0, 255).astype(np.int32)
    s, t = oracle.par_reduce_i64(a)
    assert s == int(a.astype(np.int64).sum()) and t >= 1
    u = a.astype(np.uint32)
    out, _ = oracle.par_scan_inclusive_u32(u)
    assert np.array_equal(out, np.cumsum(u, dtype=np.uint32))


# ---------------------------------------------------------------- Heterosync golden
def _closed_form(n_locs, L, shifts_of_segment):
    """Each segment of (L+1) lines of 16 words: after T shifts, line k holds the
    original line max(0, k-T); originals are storage[i] = i."""
    seg = 16 * (L + 1)
    g = np.empty(n_locs, np.float32)
    for s in range(n_locs // seg):
        T = shifts_of_segment(s)
        for k in range(L + 1):
            src = max(0, k - T)
            g[s * seg + k * 16: s * seg + k * 16 + 16] = np.arange(
                s * seg + src * 16, s * seg + src * 16 + 16, dtype=np.float32)
    return g


CASES = [
    # prim, L, num_tbs, wg, num_sm, iters, repeats
    (37, 12, 8, 32, 4, 2, 3),    # SRB barrier
    (38, 12, 8, 32, 4, 1, 2),    # local exchange: 2 shifts per iteration
    (0, 20, 6, 64, 4, 3, 2),     # tree barrier, ragged TBs per SM
    (4, 30, 5, 32, 4, 1, 2),     # non-unique mutex
    (10, 40, 8, 32, 4, 1, 1),    # non-unique semaphore (writers per SM)
    (20, 30, 6, 32, 4, 1, 2),    # unique mutex, 6 TBs on 4 SMs
    (26, 10, 8, 32, 4, 2, 2),    # unique semaphore
]


def _shifts(prim, L, num_tbs, wg, num_sm, iters, reps):
    P = wg * (L + 1)
    seg_per_block = P // (16 * (L + 1))
    per_sm = -(-num_tbs // num_sm)
    RI = iters * reps
    if prim < 4 or prim >= 37:
        lx = prim in (1, 3, 38, 40, 42, 44, 46) and per_sm > 1
        return lambda s: RI * (2 if lx else 1)
    if 4 <= prim <= 7:
        return lambda s: RI * num_tbs
    if 8 <= prim <= 19:
        nsm = min(num_tbs, num_sm)
        writers = sum(1 for sm in range(nsm) if (sm % per_sm) * nsm + sm < num_tbs)
        return lambda s: RI * writers
    if 20 <= prim <= 23:
        if num_tbs < num_sm:
            return lambda s: RI
        q, r = divmod(num_tbs, num_sm)
        return lambda s: RI * (q + (1 if s // seg_per_block < r else 0))
    return lambda s: RI


@pytest.mark.parametrize("case", CASES)
def test_hs_golden_closed_form(oracle, case):
    prim, L, num_tbs, wg, num_sm, iters, reps = case
    g = oracle.hs_golden(prim, L, num_tbs, wg, num_sm, iters, reps)
    cf = _closed_form(g.size, L, _shifts(*case))
    assert np.array_equal(g, cf)


def test_hs_golden_detects_planted_double_shift(oracle):
    # a mistake planted in the expectation (one extra shift, as an early pass
    # would cause) must be caught while total shifts < L
    prim, L, num_tbs, wg, num_sm, iters, reps = CASES[0]
    g = oracle.hs_golden(prim, L, num_tbs, wg, num_sm, iters, reps)
    planted = _closed_form(g.size, L, lambda s: iters * reps + (1 if s == 3 else 0))
    assert not np.array_equal(g, planted)


def test_prim_table_and_locs(oracle):
    assert oracle.prim_id("atomicTreeBarrUniqSRB") == 37
    assert oracle.prim_id("spinSemEBO10") == 14  # first match wins (Heterosync/main.cu:2499)
    assert oracle.prim_id("PRspinSem10") == 18
    assert oracle.prim_id("nope") == -1
    assert oracle.num_locs_mult(4, 80, 80) == 1
    assert oracle.num_locs_mult(10, 160, 80) == 2
    assert oracle.num_locs_mult(22, 160, 80) == 80
    assert oracle.num_locs_mult(37, 160, 80) == 160


def test_prim_table_matches_python_front_end(oracle, hs):
    for name, pid in hs.PRIMS:
        assert oracle.prim_id(name) == hs.prim_id(name) or name.startswith("spinSemEBO")


# ---------------------------------------------------------------- graphs
def _random_csr(n, m, seed, self_loops=False):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = self_loops | (u != v)
    pairs = np.unique(np.stack([u[keep], v[keep]], 1), axis=0)
    row = np.zeros(n + 1, np.uint32)
    np.add.at(row, pairs[:, 0] + 1, 1)
    row = np.cumsum(row, dtype=np.uint32)
    return row, pairs[:, 1].astype(np.uint32)


def _scipy_csr(row, col, data=None):
    sp = pytest.importorskip("scipy.sparse")
    n = row.size - 1
    d = np.ones(col.size) if data is None else data.astype(np.float64)
    return sp.csr_matrix((d, col.astype(np.int64), row.astype(np.int64)), shape=(n, n))


def test_bfs_special_cases(oracle):
    INF = 2**31 - 1
    # path 0->1->...->9
    row = np.arange(11, dtype=np.uint32).clip(max=9)
    row = np.array(list(range(10)) + [9], np.uint32)
    col = np.arange(1, 10, dtype=np.uint32)
    assert oracle.bfs(row, col, 0).tolist() == list(range(10))
    # star from 0, plus an isolated vertex 6
    row = np.array([0, 5, 5, 5, 5, 5, 5, 5], np.uint32)
    col = np.array([1, 2, 3, 4, 5], np.uint32)
    lv = oracle.bfs(row, col, 0)
    assert lv.tolist() == [0, 1, 1, 1, 1, 1, INF]
    assert oracle.bfs(row, col, 3).tolist() == [INF, INF, INF, 0, INF, INF, INF]


def test_bfs_sssp_vs_scipy(oracle):
    csg = pytest.importorskip("scipy.sparse.csgraph")
    INF = 2**31 - 1
    for seed in range(3):
        row, col = _random_csr(300, 1200, seed)
        w = np.random.default_rng(seed + 10).integers(1, 256, col.size).astype(np.uint32)
        for src in (0, 17, 299):
            ref = csg.shortest_path(_scipy_csr(row, col), method="D", unweighted=True, indices=src)
            lv = oracle.bfs(row, col, src)
            exp = np.where(np.isinf(ref), INF, ref).astype(np.int64)
            assert np.array_equal(lv.astype(np.int64), exp)
            ref = csg.dijkstra(_scipy_csr(row, col, w), indices=src)
            d = oracle.sssp(row, col, w, src)
            exp = np.where(np.isinf(ref), INF, ref).astype(np.int64)
            assert np.array_equal(d.astype(np.int64), exp)
        # unit weights reduce SSSP to BFS levels
        assert np.array_equal(oracle.sssp(row, col, np.ones_like(col), 5), oracle.bfs(row, col, 5))


def _reverse_csr(row, col):
    n = row.size - 1
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(row.astype(np.int64)))
    order = np.argsort(col, kind="stable")
    rrow = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n))]).astype(np.uint32)
    return rrow, src[order].astype(np.uint32)


def test_cpu_graph_baselines_vs_scipy(oracle):
    """The host baselines timed beside the GPU (direction-optimizing BFS,
    per-source Dijkstra, FP32 pull PageRank) give the pinned answers."""
    csg = pytest.importorskip("scipy.sparse.csgraph")
    INF = 2**31 - 1
    for seed in range(3):
        # average degree 15: by the second level the frontier's out-edges pass
        # 1/14 of the unexplored ones (bottom-up), and the last levels are small
        # again (top-down)
        row, col = _random_csr(2000, 30000, seed + 40)
        rrow, rcol = _reverse_csr(row, col)
        srcs = [0, 7, 1999, 7, 500]
        lv, threads = oracle.cpu_msbfs(row, col, rrow, rcol, srcs)
        assert threads >= 1
        ref = csg.shortest_path(_scipy_csr(row, col), method="D", unweighted=True, indices=srcs)
        assert np.array_equal(lv.astype(np.int64), np.where(np.isinf(ref), INF, ref).astype(np.int64))
        w = np.random.default_rng(seed).integers(1, 256, col.size).astype(np.uint32)
        d, _ = oracle.cpu_mssp(row, col, w, srcs)
        ref = csg.dijkstra(_scipy_csr(row, col, w), indices=srcs)
        assert np.array_equal(d.astype(np.int64), np.where(np.isinf(ref), INF, ref).astype(np.int64))
        x, it, _ = oracle.cpu_pagerank_pull_f32(row, rrow, rcol, alpha=0.85, tol=1e-7)
        ref, _ = oracle.pagerank(row, col, alpha=0.85, tol=1e-14)
        assert it > 2 and np.abs(x.astype(np.float64) - ref).max() < 1e-6
    # a path needs every level top-down; a star from its hub goes bottom-up at once
    n = 300
    row = np.arange(n + 1, dtype=np.uint32)
    row[-1] = n - 1
    col = np.arange(1, n, dtype=np.uint32)
    rrow, rcol = _reverse_csr(row, col)
    lv, _ = oracle.cpu_msbfs(row, col, rrow, rcol, [0, n - 1])
    assert lv[0].tolist() == list(range(n)) and lv[1].tolist() == [INF] * (n - 1) + [0]
    row = np.array([0, n - 1] + [n - 1] * (n - 1), np.uint32)
    rrow, rcol = _reverse_csr(row, col)
    lv, _ = oracle.cpu_msbfs(row, col, rrow, rcol, [0])
    assert lv[0].tolist() == [0] + [1] * (n - 1)


def test_pagerank_closed_forms(oracle):
    n = 7
    # directed cycle: uniform
    row = np.arange(n + 1, dtype=np.uint32)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)
    x, _ = oracle.pagerank(row, col)
    assert np.allclose(x, 1.0 / n, atol=1e-14)
    # complete graph without self loops: uniform
    row = (np.arange(n + 1) * (n - 1)).astype(np.uint32)
    col = np.array([v for u in range(n) for v in range(n) if v != u], np.uint32)
    x, _ = oracle.pagerank(row, col)
    assert np.allclose(x, 1.0 / n, atol=1e-14)
    # star: leaves 1..k point to hub 0, hub dangling
    k, a = 5, 0.85
    row = np.array([0] + list(range(0, k + 1)), np.uint32)
    row = np.array([0, 0] + list(range(1, k + 1)), np.uint32)
    col = np.zeros(k, np.uint32)
    x, _ = oracle.pagerank(row, col, alpha=a)
    N = k + 1
    # leaves: l = (1-a)/N + a*h/N ; hub: h = (1-a)/N + a*h/N + a*k*l
    A = np.array([[1 - a / N - a * k * 0, 0.0], [0.0, 0.0]])
    # solve the 2x2 system for (h, l)
    M = np.array([[1 - a / N, -a * k], [-a / N, 1.0]])
    rhs = np.array([(1 - a) / N, (1 - a) / N])
    h, l = np.linalg.solve(M, rhs)
    assert abs(x[0] - h) < 1e-13 and np.allclose(x[1:], l, atol=1e-13)
    assert abs(x.sum() - 1.0) < 1e-12
    del A


def test_pagerank_vs_dense_solve(oracle):
    a = 0.85
    for seed in range(2):
        row, col = _random_csr(120, 500, seed + 5)
        n = row.size - 1
        deg = np.diff(row.astype(np.int64))
        Mt = np.zeros((n, n))
        for u in range(n):
            if deg[u] == 0:
                Mt[:, u] = 1.0 / n
            for e in range(row[u], row[u + 1]):
                Mt[col[e], u] += 1.0 / deg[u]
        ref = np.linalg.solve(np.eye(n) - a * Mt, np.full(n, (1 - a) / n))
        x, _ = oracle.pagerank(row, col, alpha=a)
        assert np.abs(x - ref).max() < 1e-12
        # the reference's un-normalized push fixpoint: p = (1-a) + a * A p
        A = np.zeros((n, n))
        for u in range(n):
            for e in range(row[u], row[u + 1]):
                A[col[e], u] += 1.0 / deg[u]
        ref_p = np.linalg.solve(np.eye(n) - a * A, np.full(n, 1 - a))
        p, _ = oracle.pagerank_push_fixpoint(row, col, alpha=a)
        assert np.abs(p - ref_p).max() < 1e-11


def test_gr_fixture_and_roundtrip(oracle, tmp_path):
    row, col, w = oracle.gr_read(os.path.join(HERE, "golden", "tiny.gr"))
    assert row.tolist() == [0, 2, 3, 3] and col.tolist() == [1, 2, 2] and w.tolist() == [5, 7, 1]
    r2, c2 = _random_csr(50, 201, 9)
    w2 = np.arange(c2.size, dtype=np.uint32)
    p = str(tmp_path / "x.gr")
    assert oracle.gr_write(p, r2, c2, w2) == 0
    r3, c3, w3 = oracle.gr_read(p)
    assert np.array_equal(r2, r3) and np.array_equal(c2, c3) and np.array_equal(w2, w3)
    assert os.path.getsize(p) == 32 + 8 * 50 + 4 * c2.size + (4 if c2.size % 2 else 0) + 4 * c2.size
