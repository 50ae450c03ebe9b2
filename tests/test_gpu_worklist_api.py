"""The worklist / barrier device API driven by an application of its own.

tpds22-artifact_amd/examples/cc_labelprop.hip is built against the public
headers only (hs_worklist.hpp frontier queues and wave expansion, hs_device.hpp
grid barrier) and launched through hs_persistent_launch; its connected
components are checked against scipy's, for every barrier kind it supports and
on the edge cases of the reference's graph tests (isolated vertices, a path, a
star, an edgeless graph).  The role of a further app on LonestarGPU's
Worklist2 / PipeContextT runtime (LONESTAR/rt/include/worklist.h:49-585,
LONESTAR/rt/include/pipe.h:29-123)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "..", "tpds22-artifact_amd", "examples", "libcc_labelprop.so")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture(scope="module")
def cc_lib(hs):
    lib = C.CDLL(EXAMPLE)
    lib.hs_example_cc_work_size.argtypes = [C.c_uint32, C.POINTER(C.c_uint64)]
    lib.hs_example_cc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return lib


def symmetric_csr(n, src, dst):
    s = np.concatenate([src, dst]).astype(np.int64)
    d = np.concatenate([dst, src]).astype(np.int64)
    order = np.lexsort((d, s))
    s, d = s[order], d[order]
    row = np.zeros(n + 1, np.uint32)
    row[1:] = np.cumsum(np.bincount(s, minlength=n))
    return row, d.astype(np.uint32)


def expected_labels(n, row, col):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    a = csr_matrix((np.ones(col.size, np.int8), col.astype(np.int64), row.astype(np.int64)), shape=(n, n))
    _, comp = connected_components(a, directed=False)
    first = np.full(comp.max() + 1, n, np.int64)
    np.minimum.at(first, comp, np.arange(n))
    return first[comp].astype(np.uint32)


def run_cc(hs, torch, arena, lib, row, col, kind, num_wgs=0):
    n = row.size - 1
    d_row = torch.from_numpy(row.view(np.int32)).cuda()
    d_col = torch.from_numpy(col.view(np.int32)).cuda() if col.size else torch.zeros(1, dtype=torch.int32,
                                                                                      device="cuda")
    labels = torch.empty(n, dtype=torch.int32, device="cuda")
    nb = C.c_uint64(0)
    assert lib.hs_example_cc_work_size(n, C.byref(nb)) == 0
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    levels = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = lib.hs_example_cc(arena.handle, n, col.size, d_row.data_ptr(), d_col.data_ptr(), labels.data_ptr(),
                           work.data_ptr(), hs.barrier_kind(kind), num_wgs, levels.data_ptr(),
                           torch.cuda.current_stream().cuda_stream)
    assert rc == 0, rc
    arena.sync()
    return labels.cpu().numpy().view(np.uint32), int(levels.item())


def _graphs():
    import hsync.graph as G
    out = []
    for scale, ef, seed in ((12, 4, 3), (14, 8, 5)):
        r, c = G.rmat_csr(scale, ef, seed)
        n = r.size - 1
        src = np.repeat(np.arange(n), np.diff(r.astype(np.int64)))
        out.append((f"rmat{scale}", n, src, c.astype(np.int64)))
    n = 1000
    out.append(("path", n, np.arange(n - 1), np.arange(1, n)))
    out.append(("star_and_isolated", 300, np.zeros(199, np.int64), np.arange(1, 200)))
    out.append(("edgeless", 17, np.zeros(0, np.int64), np.zeros(0, np.int64)))
    # two cliques joined through their largest ids, and isolated vertices between
    a = np.array([(i, j) for i in range(20) for j in range(i + 1, 20)])
    b = a + 40
    out.append(("cliques", 80, np.concatenate([a[:, 0], b[:, 0], [19]]), np.concatenate([a[:, 1], b[:, 1], [59]])))
    return out


@pytest.mark.parametrize("kind", ["xcd", "flat", "two_level", "coop"])
def test_cc_through_public_api(hs, torch, arena, cc_lib, kind):
    for name, n, src, dst in _graphs():
        print("cc", kind, name, flush=True)
        row, col = symmetric_csr(n, src, dst)
        got, levels = run_cc(hs, torch, arena, cc_lib, row, col, kind)
        assert np.array_equal(got, expected_labels(n, row, col)), (kind, name)
        assert levels >= 1


def test_cc_full_grid_and_path_depth(hs, torch, arena, cc_lib):
    """8 workgroups per CU (every wave slot) on the XCD barrier; on a
    1000-vertex path the smallest label moves at least one hop per level (more
    when a neighbour in the same frontier reads an already lowered label), so
    the labels settle within 1000 levels."""
    row, col = symmetric_csr(1000, np.arange(999), np.arange(1, 1000))
    got, levels = run_cc(hs, torch, arena, cc_lib, row, col, "xcd", num_wgs=8 * 256)
    assert np.array_equal(got, np.zeros(1000, np.uint32))
    assert 1 <= levels <= 1000
