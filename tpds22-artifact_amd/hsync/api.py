"""Host-side Python API over libhs.

Device memory and streams come from PyTorch (plumbing only); every computation
runs in libhs's hand-written HIP kernels through the C ABI.  Functions mirror
the reference's host interfaces: ``reduce`` (Reduce/reduce.cuh:19),
``sum_scan_blelloch`` (scan/scan.h:18), and the Heterosync harness
(``run_sync_prim``, Heterosync/main.cu:2372-3180).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

from . import _lib as L

# Heterosync syncPrim names -> ids (Heterosync/main.cu:2481-2532).  The first
# match wins, so "spinSemEBO*" resolves to 12-15 as in the reference.
PRIMS = [
    ("atomicTreeBarrUniq", 0), ("atomicTreeBarrUniqLocalExch", 1), ("lfTreeBarrUniq", 2),
    ("lfTreeBarrUniqLocalExch", 3), ("spinMutex", 4), ("spinMutexEBO", 5), ("sleepMutex", 6),
    ("faMutex", 7), ("spinSem1", 8), ("spinSem2", 9), ("spinSem10", 10), ("spinSem120", 11),
    ("spinSemEBO1", 12), ("spinSemEBO2", 13), ("spinSemEBO10", 14), ("spinSemEBO120", 15),
    ("PRspinSem1", 16), ("PRspinSem2", 17), ("PRspinSem10", 18), ("PRspinSem120", 19),
    ("spinMutexUniq", 20), ("spinMutexEBOUniq", 21), ("sleepMutexUniq", 22), ("faMutexUniq", 23),
    ("spinSemUniq1", 24), ("spinSemUniq2", 25), ("spinSemUniq10", 26), ("spinSemUniq120", 27),
    ("spinSemEBOUniq1", 28), ("spinSemEBOUniq2", 29), ("spinSemEBOUniq10", 30),
    ("spinSemEBOUniq120", 31), ("spinSemEBO1", 32), ("spinSemEBO2", 33), ("spinSemEBO10", 34),
    ("spinSemEBO120", 35), ("atomicTreeBarrUniqSRB", 37), ("atomicTreeBarrUniqLocalExchSRB", 38),
    ("atomicTreeBarrUniqNaiveSRB", 39), ("atomicTreeBarrUniqLocalExchNaiveSRB", 40),
    ("atomicTreeBarrUniqNaiveAllSRB", 41), ("atomicTreeBarrUniqLocalExchNaiveAllSRB", 42),
    ("atomicTreeBarrUniqHybridSRB", 43), ("atomicTreeBarrUniqLocalExchHybridSRB", 44),
    ("atomicTreeBarrUniqCCG", 45), ("atomicTreeBarrUniqLocalExchCCG", 46),
]


def prim_id(name: str) -> int:
    for n, i in PRIMS:
        if n == name:
            return i
    raise ValueError(f"ERROR: Unknown synchronization primitive: {name}")


def num_locs_mult(prim: int, num_tbs: int, num_sm: int) -> int:
    """numLocsMult (Heterosync/main.cu:2558-2579)."""
    per_sm = -(-num_tbs // num_sm)
    if prim < 4 or prim >= 37 or 24 <= prim <= 35:
        return num_tbs
    if 4 <= prim <= 7:
        return 1
    if 8 <= prim <= 19:
        return per_sm
    if 20 <= prim <= 23 or prim == 36:
        return min(num_tbs, num_sm)
    raise ValueError(prim)


def storage_locs(prim: int, num_ldst: int, num_tbs: int, wg_size: int, num_sm: int) -> int:
    return num_locs_mult(prim, num_tbs, num_sm) * wg_size * (num_ldst + 1)


def _torch():
    import torch
    return torch


def _stream(stream=None) -> int:
    if stream is not None:
        return int(stream)
    torch = _torch()
    return int(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return int(t.data_ptr())


def barrier_kind(kind) -> int:
    if isinstance(kind, str):
        return L.BARRIER_NAMES[kind]
    return int(kind)


def device_info() -> dict:
    lib = L.load()
    info = L.DeviceInfo()
    L.check(lib.hs_device_info_get(C.byref(info)), "hs_device_info_get")
    return {"num_cus": info.num_cus, "num_xcds": info.num_xcds,
            "max_threads_per_block": info.max_threads_per_block, "lds_per_cu": info.lds_per_cu,
            "warp_size": info.warp_size, "coop_launch": info.coop_launch,
            "hbm_bytes": info.hbm_bytes, "arch": info.arch.decode()}


def max_resident_wgs(kind, wg_size: int) -> int:
    lib = L.load()
    out = C.c_int(0)
    L.check(lib.hs_max_resident_wgs(barrier_kind(kind), wg_size, C.byref(out)), "hs_max_resident_wgs")
    return out.value


def set_xcd_groups(groups: int) -> None:
    """Groups of the XCD barriers for later launches (0 = one per XCD)."""
    L.check(L.load().hs_set_xcd_groups(groups), "hs_set_xcd_groups")


def probe_xcc_ids(num_wgs: int, wg_size: int = 64):
    lib = L.load()
    arr = (C.c_int * num_wgs)()
    L.check(lib.hs_probe_xcc_ids(num_wgs, wg_size, arr), "hs_probe_xcc_ids")
    return list(arr)


class Arena:
    """The lock arena (replaces cudaLocksInit/cudaLocksDestroy, Heterosync/cudaLocks.cu:3-102)."""

    def __init__(self, max_blocks: int = 0, num_mutexes: int = 8, num_semaphores: int = 24,
                 num_groups: int = 0, page_align: bool = False, spin_limit: int = 0):
        lib = L.load()
        if num_groups <= 0 or max_blocks <= 0:
            info = device_info()
            num_groups = num_groups if num_groups > 0 else info["num_cus"]
            max_blocks = max_blocks if max_blocks > 0 else 8 * info["num_cus"]
        d = L.ArenaDesc(max_blocks, num_mutexes, num_semaphores, num_groups, int(page_align),
                        spin_limit)
        h = C.c_void_p()
        L.check(lib.hs_arena_create(C.byref(d), C.byref(h)), "hs_arena_create")
        self.handle = h
        self.max_blocks = max_blocks
        self.num_groups = num_groups
        self._lib = lib

    def sync(self, stream=None) -> None:
        L.check(self._lib.hs_arena_sync(self.handle, _stream(stream)), "device wait")

    def reset(self, stream=None) -> None:
        L.check(self._lib.hs_arena_reset(self.handle, _stream(stream)), "hs_arena_reset")

    def set_abort(self, value: bool) -> None:
        L.check(self._lib.hs_arena_set_abort(self.handle, int(value)), "hs_arena_set_abort")

    def wait_value(self, word, value: int, stream=None) -> None:
        """Enqueue a bounded device wait until the uint32 `word` (a 1-element
        device tensor) equals `value`; the outcome is reported by sync()."""
        L.check(self._lib.hs_wait_value_u32(self.handle, _ptr(word), value & 0xFFFFFFFF,
                                            _stream(stream)), "hs_wait_value_u32")

    def device_status(self) -> int:
        """The device status word (0, 4 aborted, 5 timeout) without resetting the arena."""
        st = C.c_int(0)
        L.check(self._lib.hs_arena_device_status(self.handle, C.byref(st)), "hs_arena_device_status")
        return st.value

    def episode_slots(self, count: int):
        """Last episode each of the first `count` workgroups published before
        arriving at a checked barrier (read before sync(), which clears them
        after a failed wait)."""
        arr = (C.c_uint32 * count)()
        L.check(self._lib.hs_arena_episode_slots(self.handle, arr, count), "hs_arena_episode_slots")
        return list(arr)

    def sem_create(self, kind: int, number: int, count: int) -> int:
        h = C.c_uint(0)
        L.check(self._lib.hs_sem_create(self.handle, kind, number, count, C.byref(h)), "hs_sem_create")
        return h.value

    def mutex_create(self, kind: int, number: int) -> int:
        h = C.c_uint(0)
        L.check(self._lib.hs_mutex_create(self.handle, kind, number, C.byref(h)), "hs_mutex_create")
        return h.value

    def close(self) -> None:
        if self.handle:
            self._lib.hs_arena_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# persistent kernels

def reduce_i32(arena: Arena, x, barrier="xcd", num_wgs: int = 0, out=None, partials=None,
               stream=None, sync: bool = True):
    """Sum of an int32 device tensor with int (wrapping) arithmetic; returns the
    int when sync is true, else the 1-element output tensor."""
    torch = _torch()
    lib = L.load()
    n = x.numel()
    nw = C.c_uint64(0)
    L.check(lib.hs_reduce_partials_size(num_wgs, n, C.byref(nw)))
    if partials is None:
        partials = torch.empty(max(nw.value, 1), dtype=torch.int32, device=x.device)
    if out is None:
        out = torch.empty(1, dtype=torch.int32, device=x.device)
    s = _stream(stream)
    L.check(lib.hs_reduce_i32(arena.handle, _ptr(x), n, _ptr(partials), _ptr(out),
                              barrier_kind(barrier), num_wgs, s), "hs_reduce_i32")
    if not sync:
        return out
    arena.sync(s)
    return int(out.item())


def reduce_i32_stamped(arena: Arena, x, stamps, barrier="xcd", num_wgs: int = 0, out=None, partials=None,
                       stream=None):
    """hs_reduce_i32_stamped: the reduce with per-workgroup clock stamps written
    to `stamps` (u64 device tensor of >= 4 x num_wgs); enqueues only."""
    torch = _torch()
    lib = L.load()
    n = x.numel()
    nw = C.c_uint64(0)
    L.check(lib.hs_reduce_partials_size(num_wgs, n, C.byref(nw)))
    if partials is None:
        partials = torch.empty(max(nw.value, 1), dtype=torch.int32, device=x.device)
    if out is None:
        out = torch.empty(1, dtype=torch.int32, device=x.device)
    L.check(lib.hs_reduce_i32_stamped(arena.handle, _ptr(x), n, _ptr(partials), _ptr(out), barrier_kind(barrier),
                                      num_wgs, _ptr(stamps), _stream(stream)), "hs_reduce_i32_stamped")
    return out


def stamp_breakdown(stamps, num_wgs: int) -> dict:
    """Microseconds of one stamped reduce launch (100 MHz stamps): the spread of
    the workgroups' starts, first start to last loads summed (streaming), last
    loads summed to the fold's end, the whole span, and the mean streaming time
    of one workgroup."""
    import numpy as np
    t = stamps[:4 * num_wgs].cpu().numpy().astype(np.int64).reshape(num_wgs, 4)
    start, loaded, leave = t[:, 0], t[:, 1], t[:, 2]
    us = lambda ticks: round(float(ticks) / 100.0, 3)  # noqa: E731
    return {"dispatch_spread_us": us(start.max() - start.min()), "stream_us": us(loaded.max() - start.min()),
            "fold_us": us(leave.max() - loaded.max()), "span_us": us(leave.max() - start.min()),
            "wg_stream_mean_us": us((loaded - start).mean())}


def reduce(arr, N: int, threads_per_block: int = 256, arena: Optional[Arena] = None) -> int:
    """reduce() of Reduce/reduce.cuh:19: sum of the first N ints of a host
    array.  threads_per_block is accepted for interface parity; the kernel runs
    256-thread workgroups."""
    torch = _torch()
    t = torch.as_tensor(arr[:N], dtype=torch.int32).cuda()
    own = arena is None
    arena = arena or Arena()
    try:
        return reduce_i32(arena, t)
    finally:
        if own:
            arena.close()


def scan_u32(arena: Arena, x, inclusive: bool = True, out=None, barrier="xcd", num_wgs: int = 0,
             scratch=None, stream=None, sync: bool = True):
    torch = _torch()
    lib = L.load()
    n = x.numel()
    if out is None:
        out = torch.empty_like(x)
    nw = C.c_uint64(0)
    L.check(lib.hs_scan_scratch_size(num_wgs, n, C.byref(nw)))
    if scratch is None:
        scratch = torch.empty(max(nw.value, 1), dtype=torch.int32, device=x.device)
    s = _stream(stream)
    L.check(lib.hs_scan_u32(arena.handle, _ptr(x), _ptr(out), n, int(inclusive), _ptr(scratch),
                            barrier_kind(barrier), num_wgs, s), "hs_scan_u32")
    if sync:
        arena.sync(s)
    return out


def scan_u32_noalloc(arena: Arena, x, out, inclusive: bool = True, barrier="xcd", num_wgs: int = 0,
                     stream=None, sync: bool = True):
    """hs_scan_u32_noalloc: the same scan with the library's static scratch
    (the caller passes input and output; nothing is allocated)."""
    lib = L.load()
    s = _stream(stream)
    L.check(lib.hs_scan_u32_noalloc(arena.handle, _ptr(x), _ptr(out), x.numel(), int(inclusive),
                                    barrier_kind(barrier), num_wgs, s), "hs_scan_u32_noalloc")
    if sync:
        arena.sync(s)
    return out


def sum_scan_blelloch(d_out, d_in, numElems: int, arena: Optional[Arena] = None,
                      inclusive: bool = False):
    """sum_scan_blelloch of scan/scan.h:18 (exclusive by default, as the reference)."""
    own = arena is None
    arena = arena or Arena()
    try:
        scan_u32(arena, d_in[:numElems], inclusive=inclusive, out=d_out[:numElems])
    finally:
        if own:
            arena.close()
    return d_out


def _batch(scan: bool, x, offsets, out, inclusive, num_wgs, work, stream, sync):
    import numpy as np
    torch = _torch()
    lib = L.load()
    h_off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    nseg = h_off.size - 1
    d_off = torch.from_numpy(h_off.view(np.int64)).to(x.device)
    nb = C.c_uint64(0)
    L.check(lib.hs_batch_work_size(h_off.ctypes.data, nseg, C.byref(nb)), "hs_batch_work_size")
    if work is None:
        work = torch.empty(max(nb.value, 8), dtype=torch.uint8, device=x.device)
    s = _stream(stream)
    if scan:
        out = out if out is not None else torch.empty_like(x)
        L.check(lib.hs_scan_batch_u32(_ptr(x), _ptr(out), _ptr(d_off), h_off.ctypes.data, nseg,
                                      int(inclusive), _ptr(work), num_wgs, s), "hs_scan_batch_u32")
    else:
        out = out if out is not None else torch.empty(nseg, dtype=torch.int32, device=x.device)
        L.check(lib.hs_reduce_batch_i32(_ptr(x), _ptr(out), _ptr(d_off), h_off.ctypes.data, nseg,
                                        _ptr(work), num_wgs, s), "hs_reduce_batch_i32")
    if sync:
        L.check(lib.hs_batch_status(_ptr(work), s), "batched launch")
    return out


def scan_batch_u32(x, offsets, inclusive=True, out=None, num_wgs=0, work=None, stream=None, sync=True):
    """Segmented scan: segment s is x[offsets[s]:offsets[s+1]], all in one launch."""
    return _batch(True, x, offsets, out, inclusive, num_wgs, work, stream, sync)


def reduce_batch_i32(x, offsets, out=None, num_wgs=0, work=None, stream=None, sync=True):
    """Per-segment int32 sums (wrapping) of x[offsets[s]:offsets[s+1]], one launch."""
    return _batch(False, x, offsets, out, False, num_wgs, work, stream, sync)


def set_scan_round_bytes(nbytes: int) -> None:
    L.check(L.load().hs_scan_set_round_bytes(nbytes), "hs_scan_set_round_bytes")


def get_scan_round_bytes() -> int:
    v = C.c_uint64(0)
    L.check(L.load().hs_scan_get_round_bytes(C.byref(v)), "hs_scan_get_round_bytes")
    return int(v.value)


class TuneResult(C.Structure):
    """hs_tune_result (include/hs.h)"""
    _fields_ = [("algo", C.c_int), ("num_wgs", C.c_int), ("barrier_kind", C.c_int),
                ("round_bytes", C.c_uint64), ("ms", C.c_float), ("default_ms", C.c_float),
                ("tried", C.c_int), ("xcd_groups", C.c_int)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def tune_reduce_i32(arena: Arena, x, reps: int = 10, stream=None) -> dict:
    """native autotuner (hs_tune_reduce_i32): the fastest grid / barrier kind"""
    r = TuneResult()
    L.check(L.load().hs_tune_reduce_i32(arena.handle, _ptr(x), x.numel(), reps, C.byref(r),
                                        _stream(stream)), "hs_tune_reduce_i32")
    return r.as_dict()


def tune_barrier(arena: Arena, storage, num_wgs: int, wg_size: int = 256, num_ldst: int = 16,
                 episodes: int = 10000, reps: int = 2, stream=None) -> dict:
    """native autotuner (hs_tune_barrier): the XCD barrier's fan-in (groups) on
    this barrier workload; apply result["xcd_groups"] with set_xcd_groups"""
    r = TuneResult()
    L.check(L.load().hs_tune_barrier(arena.handle, _ptr(storage), num_wgs, wg_size, num_ldst, episodes, reps,
                                     C.byref(r), _stream(stream)), "hs_tune_barrier")
    return r.as_dict()


def tune_scan_u32(arena: Arena, x, out, inclusive: bool = True, reps: int = 3, stream=None) -> dict:
    """native autotuner (hs_tune_scan_u32): barrier scan grid / round size or look-back"""
    r = TuneResult()
    L.check(L.load().hs_tune_scan_u32(arena.handle, _ptr(x), _ptr(out), x.numel(), int(inclusive), reps,
                                      C.byref(r), _stream(stream)), "hs_tune_scan_u32")
    return r.as_dict()


def tune_scan_batch_u32(x, offsets, out, inclusive: bool = True, reps: int = 3, stream=None) -> dict:
    """native autotuner (hs_tune_scan_batch_u32): the grid of the batched look-back scan"""
    import numpy as np
    torch = _torch()
    h_off = np.ascontiguousarray(offsets, dtype=np.uint64)
    d_off = torch.from_numpy(h_off.view(np.int64)).to(x.device)
    r = TuneResult()
    L.check(L.load().hs_tune_scan_batch_u32(_ptr(x), _ptr(out), _ptr(d_off), h_off.ctypes.data, h_off.size - 1,
                                            int(inclusive), reps, C.byref(r), _stream(stream)),
            "hs_tune_scan_batch_u32")
    return r.as_dict()


def gen_uniform_u32(out, seed: int, lo: int, hi: int, stream=None) -> None:
    L.check(L.load().hs_gen_uniform_u32(_ptr(out), out.numel(), seed, lo, hi, _stream(stream)),
            "hs_gen_uniform_u32")


def gen_uniform_u32_host(n: int, seed: int, lo: int, hi: int):
    import numpy as np
    a = np.empty(n, dtype=np.uint32)
    L.check(L.load().hs_gen_uniform_u32_host(a.ctypes.data, n, seed, lo, hi), "hs_gen_uniform_u32_host")
    return a


# ---------------------------------------------------------------------------
# microbenchmarks

def check_words(num_groups: int) -> int:
    """size of the invariant-counter buffer: violations + one padded counter per group"""
    return 64 * (num_groups + 1)


def barrier_bench(arena: Arena, kind, storage, num_wgs: int, wg_size: int, num_ldst: int,
                  episodes: int, check: bool = True, violations=None, stream=None,
                  sync: bool = True):
    torch = _torch()
    lib = L.load()
    if violations is None and check:
        violations = torch.zeros(check_words(arena.num_groups), dtype=torch.int32, device=storage.device)
    s = _stream(stream)
    L.check(lib.hs_barrier_bench(arena.handle, barrier_kind(kind), _ptr(storage), num_wgs, wg_size,
                                 num_ldst, episodes, int(check), _ptr(violations), s),
            "hs_barrier_bench")
    if sync:
        arena.sync(s)
    return violations


def sync_bench(arena: Arena, prim, storage, num_tbs: int, wg_size: int, num_ldst: int,
               iterations: int, check: bool = True, violations=None, stream=None,
               sync: bool = True):
    torch = _torch()
    lib = L.load()
    if isinstance(prim, str):
        prim = prim_id(prim)
    if violations is None and check:
        violations = torch.zeros(check_words(arena.num_groups), dtype=torch.int32, device=storage.device)
    s = _stream(stream)
    L.check(lib.hs_sync_bench(arena.handle, prim, _ptr(storage), num_tbs, wg_size, num_ldst,
                              iterations, int(check), _ptr(violations), s), "hs_sync_bench")
    if sync:
        arena.sync(s)
    return violations


def sem_bench(arena: Arena, kind: int, sem: int, storage, num_wgs: int, wg_size: int,
              num_ldst: int, iterations: int, writer_every: int, violations=None, stream=None,
              sync: bool = True):
    torch = _torch()
    lib = L.load()
    if violations is None:
        violations = torch.zeros(128, dtype=torch.int32, device=storage.device)
    s = _stream(stream)
    L.check(lib.hs_sem_bench(arena.handle, kind, sem, _ptr(storage), num_wgs, wg_size, num_ldst,
                             iterations, writer_every, _ptr(violations), s), "hs_sem_bench")
    if sync:
        arena.sync(s)
    return violations


def sem_wait_stats(violations, acquisitions: int) -> dict:
    """the wait figures hs_sem_bench leaves in its check buffer (microseconds)"""
    v = [int(x) & 0xFFFFFFFF for x in violations[:4].cpu().tolist()]
    total = v[2] | (v[3] << 32)
    return {"violations": v[0], "max_wait_us": round(v[1] * 0.01, 2),
            "mean_wait_us": round(total * 0.01 / max(acquisitions, 1), 2), "acquisitions": acquisitions}


def run_sync_prim(name: str, num_ldst: int, num_tbs: int, num_iters: int, wg_size: int = 64,
                  repeats: int = 10, arena: Optional[Arena] = None):
    """The Heterosync harness (Heterosync/main.cu:2372-3180) for one primitive:
    allocates storage[i] = i, launches the kernel `repeats` times (NUM_REPEATS),
    and returns (average ms per launch, storage as a host float32 array,
    violations)."""
    torch = _torch()
    prim = prim_id(name)
    own = arena is None
    arena = arena or Arena()
    try:
        G = arena.num_groups
        n = storage_locs(prim, num_ldst, num_tbs, wg_size, G)
        storage = torch.arange(n, dtype=torch.float32, device="cuda")
        viol_total = 0
        times = []
        for _ in range(repeats):
            st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            st.record()
            v = sync_bench(arena, prim, storage, num_tbs, wg_size, num_ldst, num_iters, check=True)
            en.record()
            torch.cuda.synchronize()
            times.append(st.elapsed_time(en))
            viol_total += int(v[0].item())
        return sum(times) / len(times), storage.cpu().numpy(), viol_total
    finally:
        if own:
            arena.close()
