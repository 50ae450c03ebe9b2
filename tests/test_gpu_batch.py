"""Batched (work-stealing, look-back) scans and reduces vs the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _segments(rng, nseg, lo_log, hi_log, zero=True):
    lens = np.exp2(rng.uniform(lo_log, hi_log, nseg)).astype(np.int64)
    if zero:
        lens[::7] = 0      # empty segments
        lens[1::11] = 1
        lens[2::13] = 4097  # one element past a tile
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)


@pytest.mark.parametrize("inclusive", [True, False])
def test_scan_batch_vs_oracle(hs, oracle, inclusive):
    import torch
    rng = np.random.default_rng(4)
    off = _segments(rng, 300, 0, 15)
    n = int(off[-1])
    x = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")[:n]
    hs.gen_uniform_u32(x, 21, 0, 2**32 - 1)
    host = oracle.gen_uniform_u32(21, n, 0, 2**32 - 1)
    out = hs.scan_batch_u32(x, off, inclusive=inclusive).cpu().numpy().view(np.uint32)
    for s in range(off.size - 1):
        a, b = int(off[s]), int(off[s + 1])
        assert np.array_equal(out[a:b], oracle.scan_u32(host[a:b], inclusive)), s


def test_scan_batch_unaligned_and_grids(hs, oracle):
    import torch
    off = np.array([0, 3, 5000, 5001, 9000, 40000, 40001, 100003], np.uint64)  # unaligned starts
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 5, 0, 15)
    host = oracle.gen_uniform_u32(5, n, 0, 15)
    for wgs in (0, 1, 7, 64):
        out = hs.scan_batch_u32(x, off, num_wgs=wgs).cpu().numpy().view(np.uint32)
        for s in range(off.size - 1):
            a, b = int(off[s]), int(off[s + 1])
            assert np.array_equal(out[a:b], oracle.scan_u32(host[a:b], True)), (wgs, s)


def test_reduce_batch_vs_oracle(hs, oracle):
    import torch
    rng = np.random.default_rng(8)
    off = _segments(rng, 500, 0, 16)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 3, 0, 2**32 - 1)
    host = oracle.gen_uniform_u32(3, n, 0, 2**32 - 1).view(np.int32)
    sums = hs.reduce_batch_i32(x, off).cpu().numpy()
    for s in range(off.size - 1):
        a, b = int(off[s]), int(off[s + 1])
        assert int(sums[s]) == oracle.reduce_i32(host[a:b]), s


def test_scan_batch_headline_shape(hs, oracle):
    """512 scans of log-uniform lengths 2^10..2^22 in one launch; every segment
    against the oracle, and every segment's last value against its sum."""
    import torch
    rng = np.random.default_rng(0)
    lens = np.exp2(rng.uniform(10, 22, 512)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    hs.gen_uniform_u32(x, 0, 0, 15)
    out = hs.scan_batch_u32(x, off)
    host = oracle.gen_uniform_u32(0, n, 0, 15)
    o = out.cpu().numpy().view(np.uint32)
    for s in range(512):
        a, b = int(off[s]), int(off[s + 1])
        assert np.array_equal(o[a:b], oracle.scan_u32(host[a:b], True)), s
    ends = off[1:].astype(np.int64) - 1
    seg_sums = np.add.reduceat(host.astype(np.uint64), off[:-1].astype(np.int64)) % 2**32
    assert np.array_equal(o[ends].astype(np.uint64), seg_sums)
