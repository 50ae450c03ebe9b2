"""hs_scan_u32_noalloc (the scan with the library's static scratch) against
hs_scan_u32 with a caller-allocated scratch and against the oracle, through the
C ABI: two calls in a row, the second smaller, then the refusals."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_scan_noalloc_matches_scan(hs, arena, oracle):
    import torch
    cases = [((3 << 20) + 123, 0, "xcd", True, 21),   # default grid (2 per CU), 4 MiB rounds
             (100003, 96, "two_level", False, 22)]     # smaller input, smaller grid, exclusive
    hs.set_scan_round_bytes(4 << 20)  # several rounds: both partial buffers are used
    try:
        for n, wgs, kind, inclusive, seed in cases:
            print("noalloc scan", n, wgs, kind, inclusive, flush=True)
            x = torch.empty(n, dtype=torch.int32, device="cuda")
            hs.gen_uniform_u32(x, seed, 0, 2**32 - 1)
            ref = hs.scan_u32(arena, x, inclusive=inclusive, barrier=kind, num_wgs=wgs)
            out = torch.full_like(x, -1)
            hs.scan_u32_noalloc(arena, x, out, inclusive=inclusive, barrier=kind, num_wgs=wgs)
            assert torch.equal(out, ref), (n, wgs, kind)
            host = oracle.gen_uniform_u32(seed, n, 0, 2**32 - 1)
            assert np.array_equal(out.cpu().numpy().view(np.uint32), oracle.scan_u32(host, inclusive))
    finally:
        hs.set_scan_round_bytes(64 << 20)


def test_scan_noalloc_refusals(hs, arena):
    import torch
    x = torch.zeros(4096, dtype=torch.int32, device="cuda")
    out = torch.empty_like(x)
    for kind, wgs in (("xcd", 4097), ("coop", 0)):
        with pytest.raises(hs.HSError) as e:
            hs.scan_u32_noalloc(arena, x, out, barrier=kind, num_wgs=wgs)
        assert e.value.status == 1
