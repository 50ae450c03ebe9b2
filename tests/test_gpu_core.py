"""GPU parity tests: barriers, reduce, scan and the Heterosync primitives
against the CPU oracle, through the C ABI (libhs.so)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ["flat", "two_level", "xcd", "xcd_flat_release", "tree", "lf", "coop", "naive_all"]


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def test_device_info_and_xcd_dispatch(hs, torch):
    info = hs.device_info()
    _log("device", info)
    assert info["arch"].startswith("gfx950")
    assert info["warp_size"] == 64 and info["num_cus"] >= 8
    ids = hs.probe_xcc_ids(4 * info["num_cus"], 64)
    assert max(ids) + 1 == info["num_xcds"]
    rr = sum(1 for i, x in enumerate(ids) if x == i % info["num_xcds"]) / len(ids)
    _log("fraction of WGs on XCD (wg % nxcd):", rr)


# ------------------------------------------------------------------ reduce
@pytest.mark.parametrize("n", [0, 1, 3, 5, 4096, 1000003, 1 << 22])
def test_reduce_sizes(hs, arena, oracle, torch, n):
    _log("reduce n", n)
    x = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")[:n]
    hs.gen_uniform_u32(x.view(torch.int32), 5, 0, 255) if n else None
    host = oracle.gen_uniform_u32(5, n, 0, 255).astype(np.int32)
    if n:
        assert np.array_equal(x.cpu().numpy(), host)  # device generator == oracle generator
    assert hs.reduce_i32(arena, x) == oracle.reduce_i32(host)


@pytest.mark.parametrize("kind", KINDS)
def test_reduce_all_barriers(hs, arena, oracle, torch, kind):
    _log("reduce kind", kind)
    n = (1 << 20) + 7
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 11, 0, 2**32 - 1)  # full range: exercises int32 wrap
    exp = oracle.reduce_i32(oracle.gen_uniform_u32(11, n, 0, 2**32 - 1).view(np.int32))
    for wgs in (0, 64, 1024):
        assert hs.reduce_i32(arena, x, barrier=kind, num_wgs=wgs) == exp


def test_reduce_full_size(hs, arena, oracle, torch):
    n = 1 << 24  # the headline workload: 2^24 int32 uniform [0,255]
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 0, 0, 255)
    host = oracle.gen_uniform_u32(0, n, 0, 255).astype(np.int32)
    exp64 = oracle.reduce_i64(host)
    assert exp64 < 2**31 - 1  # no overflow at this seed
    for kind in ("xcd", "two_level", "flat", "coop"):
        assert hs.reduce_i32(arena, x, barrier=kind) == exp64


def test_reduce_ones_reference_case(hs, arena, torch):
    # Reduce/test_reduce.cu:14-17: all-ones input, correct_sum = N
    for n in (2560, 81920):
        assert hs.reduce(np.ones(n, np.int32), n, 32, arena=arena) == n


# ------------------------------------------------------------------ scan
@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 1000003])
@pytest.mark.parametrize("inclusive", [True, False])
def test_scan_sizes(hs, arena, oracle, torch, n, inclusive):
    _log("scan", n, inclusive)
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 3, 0, 2**32 - 1)
    host = oracle.gen_uniform_u32(3, n, 0, 2**32 - 1)
    out = hs.scan_u32(arena, x, inclusive=inclusive)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), oracle.scan_u32(host, inclusive))


@pytest.mark.parametrize("kind", KINDS)
def test_scan_all_barriers_multi_round(hs, arena, oracle, torch, kind):
    _log("scan kind", kind)
    n = 3 * (1 << 20) + 123
    hs.set_scan_round_bytes(1 << 20)  # force many rounds
    try:
        x = torch.empty(n, dtype=torch.int32, device="cuda")
        hs.gen_uniform_u32(x, 9, 0, 15)
        host = oracle.gen_uniform_u32(9, n, 0, 15)
        for wgs in (0, 96):
            out = hs.scan_u32(arena, x, inclusive=True, barrier=kind, num_wgs=wgs)
            assert np.array_equal(out.cpu().numpy().view(np.uint32), oracle.scan_u32(host, True))
    finally:
        hs.set_scan_round_bytes(96 << 20)


def test_scan_reference_case_ones(hs, arena, torch):
    # scan/main.cu:36-41 input (all ones), exclusive: out[i] = i
    n = 163840
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    d_out = torch.empty_like(x)
    hs.sum_scan_blelloch(d_out, x, n, arena=arena)
    assert torch.equal(d_out.cpu(), torch.arange(n, dtype=torch.int32))


def test_scan_full_size_properties(hs, arena, torch):
    n = 1 << 28  # headline: 2^28 int32 uniform [0,15], inclusive
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 0, 0, 15)
    out = hs.scan_u32(arena, x, inclusive=True)
    total = int(x.sum(dtype=torch.int64).item())
    assert total < 2**31
    assert int(out[-1].item()) == total
    assert int(out[0].item()) == int(x[0].item())
    # differences reproduce the input (size-independent property); chunked
    step = 1 << 26
    for b in range(0, n, step):
        e = min(n, b + step)
        lo = out[b - 1:e - 1] if b else torch.cat([out.new_zeros(1), out[:e - 1]])
        assert torch.equal(out[b:e] - lo, x[b:e])


# ------------------------------------------------------------------ barriers
def _golden(oracle, prim, L, tbs, wg, nsm, iters, reps):
    return oracle.hs_golden(prim, L, tbs, wg, nsm, iters, reps)


@pytest.mark.parametrize("kind", KINDS)
def test_barrier_invariants_and_golden(hs, arena, oracle, torch, kind):
    _log("barrier", kind)
    wg, L, iters = 64, 24, 5  # iters < L: an early pass changes the result
    for tbs in (7, 64, 600):
        n = tbs * wg * (L + 1)
        storage = torch.arange(n, dtype=torch.float32, device="cuda")
        v = hs.barrier_bench(arena, kind, storage, tbs, wg, L, iters, check=True)
        assert int(v[0].item()) == 0
        g = _golden(oracle, 37, L, tbs, wg, arena.num_groups, iters, 1)
        assert np.array_equal(storage.cpu().numpy(), g), (kind, tbs)


@pytest.mark.parametrize("kind", ["flat", "two_level", "xcd", "xcd_flat_release", "lf", "coop"])
def test_barrier_headline_shape_no_early_pass(hs, arena, torch, kind):
    # 1024 WGs x 256 threads, 16 stores per thread per episode
    wgs, wg, L, eps = 1024, 256, 16, 20000
    storage = torch.zeros(wgs * wg * (L + 1), dtype=torch.float32, device="cuda")
    v = hs.barrier_bench(arena, kind, storage, wgs, wg, L, eps, check=True)
    assert int(v[0].item()) == 0


def _run_checked(hs, arena, torch, kind, storage, wgs, wg, L, eps):
    """One checked barrier launch; on a failed wait, report which workgroups
    had arrived at which episode (slot 0: never arrived) before the arena is
    reset."""
    v = hs.barrier_bench(arena, kind, storage, wgs, wg, L, eps, check=True, sync=False)
    torch.cuda.synchronize()
    st = arena.device_status()
    if st:
        slots = np.array(arena.episode_slots(wgs))
        vals, counts = np.unique(slots, return_counts=True)
        low = np.flatnonzero(slots == slots.min())[:16].tolist()
        _log(f"{kind} {wgs}x{wg}: status {st}; episode slots {dict(zip(vals.tolist(), counts.tolist()))}; "
             f"lowest at workgroups {low}")
    arena.sync()
    return int(v[0].item())


@pytest.mark.parametrize("kind", ["xcd", "flat", "two_level", "coop"])
def test_barrier_every_wave_slot_2048(hs, arena, oracle, torch, kind):
    # 2048 WGs x 256 threads: 8 per CU, every wave slot the occupancy API reports
    wgs, wg = 2048, 256
    assert hs.max_resident_wgs(kind, wg) >= wgs
    _log("barrier 2048", kind)
    L, iters = 4, 3  # golden: iters < L, so an early pass changes the result
    storage = torch.arange(wgs * wg * (L + 1), dtype=torch.float32, device="cuda")
    assert _run_checked(hs, arena, torch, kind, storage, wgs, wg, L, iters) == 0
    g = _golden(oracle, 37, L, wgs, wg, arena.num_groups, iters, 1)
    assert np.array_equal(storage.cpu().numpy(), g), kind
    # the microbench shape (16 stores per thread), invariant counters on
    storage = torch.zeros(wgs * wg * 17, dtype=torch.float32, device="cuda")
    assert _run_checked(hs, arena, torch, kind, storage, wgs, wg, 16, 2000) == 0


@pytest.mark.parametrize("groups", [1, 3, 4, 16])
def test_xcd_barrier_any_group_count(hs, arena, oracle, torch, groups):
    # groups that span several XCDs take the write-back-request path every
    # episode (3, 4, 1 groups on 8 XCDs); 16 groups stay XCD-local
    _log("xcd groups", groups)
    hs.set_xcd_groups(groups)
    try:
        wg, L, iters = 64, 24, 5
        for kind in ("xcd", "xcd_flat_release"):
            for tbs in (7, 64, 600):
                n = tbs * wg * (L + 1)
                storage = torch.arange(n, dtype=torch.float32, device="cuda")
                v = hs.barrier_bench(arena, kind, storage, tbs, wg, L, iters, check=True)
                assert int(v[0].item()) == 0, (kind, tbs)
                g = _golden(oracle, 37, L, tbs, wg, arena.num_groups, iters, 1)
                assert np.array_equal(storage.cpu().numpy(), g), (kind, tbs)
        storage = torch.zeros(512 * 256 * 17, dtype=torch.float32, device="cuda")
        v = hs.barrier_bench(arena, "xcd", storage, 512, 256, 16, 5000, check=True)
        assert int(v[0].item()) == 0
        x = torch.empty((1 << 22) + 5, dtype=torch.int32, device="cuda")
        hs.gen_uniform_u32(x, groups, 0, 2**32 - 1)
        host = oracle.gen_uniform_u32(groups, x.numel(), 0, 2**32 - 1)
        assert hs.reduce_i32(arena, x, barrier="xcd") == oracle.reduce_i32(host.view(np.int32))
        y = hs.scan_u32(arena, x, inclusive=True, barrier="xcd")
        assert np.array_equal(y.cpu().numpy().view(np.uint32), oracle.scan_u32(host, True))
    finally:
        hs.set_xcd_groups(0)


def test_residency_refused(hs, arena, torch):
    lim = hs.max_resident_wgs("xcd", 1024)
    storage = torch.zeros(1 << 20, dtype=torch.float32, device="cuda")
    with pytest.raises(hs.HSError) as e:
        hs.barrier_bench(arena, "xcd", storage, lim + 1, 1024, 0, 1, check=False)
    assert e.value.status in (1, 3)


def test_timeout_is_reported_not_hung(hs, torch):
    # a tiny spin limit forces every wait to give up: the kernel still exits
    a = hs.Arena(spin_limit=2)
    try:
        storage = torch.zeros(256 * 64 * 2, dtype=torch.float32, device="cuda")
        with pytest.raises(hs.HSError) as e:
            hs.barrier_bench(a, "flat", storage, 256, 64, 1, 50, check=False)
        assert e.value.status == 5
        # a semaphore contended past the spin limit: waiters give up while others
        # hold it, so its record is left half-acquired
        sem = a.sem_create(hs.SEM_SPIN, 3, 1)
        st = torch.zeros(17 * 256, dtype=torch.float32, device="cuda")
        with pytest.raises(hs.HSError) as e:
            hs.sem_bench(a, hs.SEM_SPIN, sem, st, 256, 256, 16, 20, 10)
        assert e.value.status == 5
        # the failure reset the arena and restored the record: one writer (it
        # needs the full capacity) acquires at its first attempt, and a barrier
        # of one workgroup (no waiting) runs
        v = hs.sem_bench(a, hs.SEM_SPIN, sem, st, 1, 256, 16, 3, 1)
        assert int(v[0].item()) == 0
        hs.barrier_bench(a, "flat", storage, 1, 64, 1, 5, check=False)
    finally:
        a.close()


def test_abort_flag(hs, torch):
    # a wait on a word the host never sets: the abort flag ends it (the spin
    # limit would end it too, with status 5, if the flag were not seen)
    a = hs.Arena()
    try:
        word = torch.zeros(1, dtype=torch.int32, device="cuda")
        a.set_abort(True)
        a.wait_value(word, 1)
        with pytest.raises(hs.HSError) as e:
            a.sync()
        assert e.value.status == 4
        a.set_abort(False)
        storage = torch.zeros(128 * 64 * 2, dtype=torch.float32, device="cuda")
        hs.barrier_bench(a, "xcd", storage, 128, 64, 1, 100, check=False)
        word.fill_(7)
        a.wait_value(word, 7)
        a.sync()
    finally:
        a.close()


def test_wait_value_timeout(hs, torch):
    a = hs.Arena(spin_limit=64)
    try:
        word = torch.zeros(1, dtype=torch.int32, device="cuda")
        a.wait_value(word, 3)
        with pytest.raises(hs.HSError) as e:
            a.sync()
        assert e.value.status == 5
    finally:
        a.close()


# ------------------------------------------------------------------ Heterosync matrix
SMALL = dict(L=20, tbs=24, wg=64, iters=2, nsm=8)


@pytest.mark.parametrize("name,pid", [p for p in __import__("hsync").PRIMS
                                      if p[1] not in range(32, 36)])
def test_heterosync_prims_vs_golden(hs, oracle, torch, name, pid):
    _log("prim", name)
    L, tbs, wg, iters, nsm = SMALL["L"], SMALL["tbs"], SMALL["wg"], SMALL["iters"], SMALL["nsm"]
    a = hs.Arena(max_blocks=256, num_groups=nsm)
    try:
        n = hs.storage_locs(pid, L, tbs, wg, nsm)
        storage = torch.arange(n, dtype=torch.float32, device="cuda")
        reps = 2
        for _ in range(reps):
            v = hs.sync_bench(a, pid, storage, tbs, wg, L, iters, check=True)
            assert int(v[0].item()) == 0, name
        g = oracle.hs_golden(pid, L, tbs, wg, nsm, iters, reps)
        assert np.array_equal(storage.cpu().numpy(), g), name
    finally:
        a.close()


@pytest.mark.parametrize("pid", [32, 33, 34, 35])
def test_heterosync_ebo_priority_sems_by_id(hs, oracle, torch, pid):
    # the EBO priority semaphores (Heterosync/cudaLocksSemaphoreEBO.cu:456-930):
    # their names resolve to ids 12-15 first (Heterosync/main.cu:2497-2500 vs
    # 2518-2521), so they are reachable by id only; golden replay of
    # Heterosync/main.cu:3067-3101
    _log("prim id", pid)
    L, tbs, wg, iters, nsm = SMALL["L"], SMALL["tbs"], SMALL["wg"], SMALL["iters"], SMALL["nsm"]
    a = hs.Arena(max_blocks=256, num_groups=nsm)
    try:
        n = hs.storage_locs(pid, L, tbs, wg, nsm)
        storage = torch.arange(n, dtype=torch.float32, device="cuda")
        reps = 2
        for _ in range(reps):
            v = hs.sync_bench(a, pid, storage, tbs, wg, L, iters, check=True)
            assert int(v[0].item()) == 0, pid
        g = oracle.hs_golden(pid, L, tbs, wg, nsm, iters, reps)
        assert np.array_equal(storage.cpu().numpy(), g), pid
    finally:
        a.close()


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("cap", [1, 16, 256])
def test_semaphore_mix_invariants(hs, arena, torch, kind, cap):
    _log("sem", kind, cap)
    sem = arena.sem_create(kind, 20, cap)
    storage = torch.zeros(256 * 17 * 2, dtype=torch.float32, device="cuda")
    v = hs.sem_bench(arena, kind, sem, storage, 256, 256, 16, 20, 10)
    assert int(v[0].item()) == 0


def test_xcd_queue_semaphore_bounded_wait_1e6(hs, arena, torch):
    """The per-XCD ticket queue over 1,000,448 acquisitions (1024 WGs x 977,
    readers:writers 9:1, capacity 16): capacity and writer exclusivity hold at
    every entry, and no acquisition waits much longer than the average one: the
    XCD queues and the turns of their heads serve waiters in arrival order, so
    each waits for about one round of all contenders (measured: the longest
    wait within 2% of the mean)."""
    _log("xcd queue 1e6 acquisitions")
    sem = arena.sem_create(hs.SEM_XCD_QUEUE, 21, 16)
    storage = torch.zeros(256 * 17 * 2, dtype=torch.float32, device="cuda")
    v = torch.zeros(128, dtype=torch.int32, device="cuda")
    hs.sem_bench(arena, hs.SEM_XCD_QUEUE, sem, storage, 1024, 256, 16, 977, 10, violations=v)
    st = hs.sem_wait_stats(v, 1024 * 977)
    print("xcd queue wait", st, flush=True)
    assert st["violations"] == 0
    assert st["mean_wait_us"] > 0
    assert st["max_wait_us"] < 2 * st["mean_wait_us"], st


@pytest.mark.parametrize("kind", ["xcd", "two_level", "flat", "coop"])
def test_reduce_stamped_same_sum_and_ordered_stamps(hs, arena, oracle, torch, kind):
    """The diagnostic stamped reduce gives the plain reduce's sum, and every
    workgroup's stamps are ordered start <= loads summed <= leave."""
    n = (1 << 22) + 5
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 9, 0, 255)
    wgs = 256
    stamps = torch.zeros(4 * wgs, dtype=torch.int64, device="cuda")
    out = hs.reduce_i32_stamped(arena, x, stamps, barrier=kind, num_wgs=wgs)
    arena.sync()
    assert int(out.item()) == hs.reduce_i32(arena, x, barrier=kind, num_wgs=wgs)
    t = stamps.cpu().numpy().reshape(wgs, 4)
    assert (t[:, 0] > 0).all() and (t[:, 1] >= t[:, 0]).all() and (t[:, 2] >= t[:, 1]).all()
    br = hs.stamp_breakdown(stamps, wgs)
    print(kind, br, flush=True)
    assert br["span_us"] >= br["stream_us"] >= br["dispatch_spread_us"] >= 0
