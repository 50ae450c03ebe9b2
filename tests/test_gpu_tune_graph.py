"""hipGraph replay equals eager launches; the autotuner keeps results and is
never slower than the default configuration it measured."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replay_reduce_and_scan(hs, arena, oracle, tmp_path):
    import torch
    from hsync.graphcap import Replay
    n = (1 << 20) + 3
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    partials = torch.empty(4096, dtype=torch.int32, device="cuda")
    scratch = torch.empty(8192, dtype=torch.int32, device="cuda")
    out = torch.empty(1, dtype=torch.int32, device="cuda")

    def work():
        hs.reduce_i32(arena, x, out=out, partials=partials, sync=False)
        hs.scan_u32(arena, x, inclusive=True, out=y, scratch=scratch, sync=False)

    hs.gen_uniform_u32(x, 0, 0, 255)
    g = Replay(work)
    for seed in range(1, 5):
        hs.gen_uniform_u32(x, seed, 0, 255)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        arena.sync()
        host = oracle.gen_uniform_u32(seed, n, 0, 255)
        assert int(out.item()) == oracle.reduce_i32(host.astype(np.int32)), seed
        assert np.array_equal(y.cpu().numpy().view(np.uint32), oracle.scan_u32(host, True)), seed
        # eager run gives the same bits
        eager = hs.reduce_i32(arena, x)
        assert eager == int(out.item())


def test_graph_replay_barrier_episodes(hs, arena):
    import torch
    from hsync.graphcap import Replay
    storage = torch.zeros(512 * 256 * 9, dtype=torch.float32, device="cuda")
    v = torch.zeros(hs.check_words(arena.num_groups), dtype=torch.int32, device="cuda")
    g = Replay(lambda: hs.barrier_bench(arena, "xcd", storage, 512, 256, 8, 200, check=True,
                                        violations=v, sync=False))
    for _ in range(3):
        g.replay()
        arena.sync()
        assert int(v[0].item()) == 0


def test_tuner_reduce_and_scan(hs, arena, oracle, tmp_path):
    import torch
    from hsync.tune import Tuner
    t = Tuner(arena, cache_path=str(tmp_path / "tune.json"))
    x = torch.empty(1 << 22, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 2, 0, 255)
    cfg, rec = t.reduce(x)
    assert rec["ms"] <= rec["default_ms"] * 1.0001
    exp = oracle.reduce_i32(oracle.gen_uniform_u32(2, x.numel(), 0, 255).astype(np.int32))
    assert hs.reduce_i32(arena, x, barrier=cfg["barrier"], num_wgs=cfg["num_wgs"]) == exp
    y = torch.empty_like(x)
    cfg2, rec2 = t.scan(x, y)
    assert rec2["ms"] <= rec2["default_ms"] * 1.0001
    # cached on the second call
    cfg3, _ = Tuner(arena, cache_path=str(tmp_path / "tune.json")).reduce(x)
    assert cfg3 == cfg


def test_native_tuner_reduce_and_scan(hs, arena, oracle):
    """hs_tune_*: the default is timed first, the choice is no slower than it,
    the choice computes the same bits, and the round size is left unchanged"""
    import torch
    x = torch.empty((1 << 22) + 5, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 3, 0, 255)
    h = oracle.gen_uniform_u32(3, x.numel(), 0, 255)
    r = hs.tune_reduce_i32(arena, x, reps=5)
    print("native reduce tune", r, flush=True)
    assert r["tried"] >= 2 and r["ms"] <= r["default_ms"] * 1.0001
    got = hs.reduce_i32(arena, x, barrier=r["barrier_kind"], num_wgs=r["num_wgs"])
    assert got == oracle.reduce_i32(h.astype(np.int32))
    round0 = hs.get_scan_round_bytes()
    y = torch.empty_like(x)
    r = hs.tune_scan_u32(arena, x, y, inclusive=True, reps=2)
    print("native scan tune", r, flush=True)
    assert r["tried"] >= 2 and r["ms"] <= r["default_ms"] * 1.0001
    assert hs.get_scan_round_bytes() == round0
    y.zero_()
    if r["algo"] == 1:
        hs.scan_batch_u32(x, [0, x.numel()], inclusive=True, out=y, num_wgs=r["num_wgs"])
    else:
        hs.set_scan_round_bytes(r["round_bytes"])
        try:
            hs.scan_u32(arena, x, inclusive=True, out=y, barrier=r["barrier_kind"], num_wgs=r["num_wgs"])
        finally:
            hs.set_scan_round_bytes(round0)
    assert np.array_equal(y.cpu().numpy().view(np.uint32), oracle.scan_u32(h, True))


def test_native_tuner_fan_in_applied(hs, arena, oracle):
    """a reduce tuning that includes XCD barriers of 2x and 4x the groups: the
    choice, applied with its group count, gives the oracle's sum"""
    import torch
    x = torch.empty(1 << 22, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 8, 0, 2**32 - 1)
    exp = oracle.reduce_i32(oracle.gen_uniform_u32(8, x.numel(), 0, 2**32 - 1).view(np.int32))
    r = hs.tune_reduce_i32(arena, x, reps=5)
    assert r["tried"] >= 10 and r["ms"] <= r["default_ms"] * 1.0001
    hs.set_xcd_groups(r["xcd_groups"])
    try:
        assert hs.reduce_i32(arena, x, barrier=r["barrier_kind"], num_wgs=r["num_wgs"]) == exp
    finally:
        hs.set_xcd_groups(0)
    for g in (16, 32):  # every fan-in candidate computes the same bits
        hs.set_xcd_groups(g)
        try:
            assert hs.reduce_i32(arena, x, barrier="xcd", num_wgs=512) == exp
        finally:
            hs.set_xcd_groups(0)


def test_native_tuner_graph_launches(hs, arena, oracle):
    """hs_tune_msbfs / _bfs / _mssp / _pagerank_f32: never slower than the
    default in the same call; the chosen configuration gives the same result"""
    import torch
    import hsync.graph as G
    row, col = G.rmat_csr(14, 16, 9)
    g = G.Graph(row, col)
    gw = G.Graph(row, col, G.weights(col.size, 9), transposed=False)
    deg = np.diff(row.astype(np.int64))
    srcs = [int(v) for v in np.random.default_rng(9).choice(np.nonzero(deg > 0)[0], 64, replace=False)]

    def applied(r, fn):
        hs.set_xcd_groups(r["xcd_groups"])
        try:
            return fn(r["barrier_kind"], r["num_wgs"])
        finally:
            hs.set_xcd_groups(0)

    r = G.tune_msbfs(arena, g, srcs, reps=2)
    print("tune msbfs", r, flush=True)
    assert r["tried"] >= 4 and r["ms"] <= r["default_ms"] * 1.0001
    lv = applied(r, lambda k, w: G.msbfs(arena, g, srcs, barrier=k, num_wgs=w)).cpu().numpy()
    for i in (0, 17, 63):
        assert np.array_equal(lv[i], oracle.bfs(row, col, srcs[i])), i
    r = G.tune_bfs(arena, g, srcs[0], reps=2)
    assert r["ms"] <= r["default_ms"] * 1.0001
    lv1, _ = applied(r, lambda k, w: G.bfs(arena, g, srcs[0], barrier=k, num_wgs=w))
    assert np.array_equal(lv1.cpu().numpy(), oracle.bfs(row, col, srcs[0]))
    r = G.tune_mssp(arena, gw, srcs[:8], reps=1)
    assert r["ms"] <= r["default_ms"] * 1.0001
    d, _ = applied(r, lambda k, w: G.mssp(arena, gw, srcs[:8], barrier=k, num_wgs=w))
    assert np.array_equal(d.cpu().numpy()[3], oracle.sssp(row, col, gw.h_w, srcs[3]))
    rank0, _ = G.pagerank(arena, g)
    r = G.tune_pagerank(arena, g, reps=1)
    assert r["ms"] <= r["default_ms"] * 1.0001
    rank1, _ = applied(r, lambda k, w: G.pagerank(arena, g, barrier=k, num_wgs=w))
    assert torch.equal(rank0, rank1)  # any grid and barrier: the same bits


def test_native_tuner_scan_batch_headline(hs, oracle):
    """the 512-segment batch (lengths 2^10..2^22) in one autotuned launch"""
    import torch
    rng = np.random.default_rng(0)
    lens = np.exp2(rng.uniform(10, 22, 512)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 4, 0, 15)
    y = torch.empty_like(x)
    r = hs.tune_scan_batch_u32(x, off, y, reps=2)
    print("tune scan batch", r, flush=True)
    assert r["tried"] >= 3 and r["ms"] <= r["default_ms"] * 1.0001
    y.zero_()
    out = hs.scan_batch_u32(x, off, out=y, num_wgs=r["num_wgs"]).cpu().numpy().view(np.uint32)
    host = oracle.gen_uniform_u32(4, n, 0, 15)
    for s in range(0, 512, 7):
        a, b = int(off[s]), int(off[s + 1])
        assert np.array_equal(out[a:b], oracle.scan_u32(host[a:b], True)), s


def test_tune_barrier_fan_in(hs, arena):
    """hs_tune_barrier: the XCD barrier's groups on a barrier workload; the
    choice is never slower than the default in the same call, and the tuned
    groups pass the per-episode invariants and the golden data check."""
    import torch
    storage = torch.zeros(1024 * 256 * 17, dtype=torch.float32, device="cuda")
    r = hs.tune_barrier(arena, storage, 1024, 256, 16, episodes=2000, reps=2)
    print("tune barrier", r, flush=True)
    assert r["tried"] >= 2 and r["ms"] <= r["default_ms"]
    assert r["xcd_groups"] == 0 or r["xcd_groups"] % 8 == 0
    hs.set_xcd_groups(r["xcd_groups"])
    try:
        v = torch.zeros(hs.check_words(arena.num_groups), dtype=torch.int32, device="cuda")
        hs.barrier_bench(arena, "xcd", storage, 1024, 256, 16, 5000, check=True, violations=v)
        assert int(v[0].item()) == 0
    finally:
        hs.set_xcd_groups(0)
