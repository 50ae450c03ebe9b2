"""CPU-side tests of libhs: the C ABI loads and exports every declared symbol,
and the host-side input generators / graph formats agree with the oracle."""
import glob
import hashlib
import json
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(hs\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(hs):
    lib = hs.load()
    declared = _declared_symbols()
    assert "hs_arena_create" in declared and "hsLocksInit" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    from hsync import _lib as L
    bound = set(L.SIGNATURES) | set(L.GRAPH_SIGNATURES)
    assert declared <= bound, declared - bound
    assert lib.hs_version().decode().startswith("libhs")


def _declared_arg_counts():
    """name -> number of parameters of every hs* function in include/*.h"""
    out = {}
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(hs\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.M):
            params = " ".join(m.group(2).split())
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_tables_match_header_arity(hs):
    """every declared function is bound with the header's number of arguments
    (a new parameter added on one side only would shift every later argument)"""
    from hsync import _lib as L
    counts = _declared_arg_counts()
    assert counts.get("hs_serve_submit_pagerank") == 11 and counts.get("hs_serve_phases") == 3
    table = {**L.SIGNATURES, **L.GRAPH_SIGNATURES}
    bad = {n: (c, len(table[n][1])) for n, c in counts.items() if n in table and len(table[n][1]) != c}
    assert not bad, bad
    assert set(counts) == _declared_symbols()


def test_status_strings(hs):
    lib = hs.load()
    assert lib.hs_status_string(3).decode() == "grid exceeds co-resident workgroups"
    assert lib.hs_status_string(0).decode() == "ok"


def test_host_generator_matches_oracle(hs, oracle):
    for seed, lo, hi in ((0, 0, 255), (99, 0, 15), (2**63 + 5, 1, 255)):
        a = hs.gen_uniform_u32_host(100003, seed, lo, hi)
        b = oracle.gen_uniform_u32(seed, 100003, lo, hi)
        assert np.array_equal(a, b)


def _rmat(hs, scale, ef, seed, a=0.57, b=0.19, c=0.19):
    import ctypes as C
    lib = hs.load()
    m = C.c_uint64(0)
    assert lib.hs_rmat_csr_host(scale, ef, a, b, c, seed, None, None, C.byref(m)) == 0
    n = 1 << scale
    row = np.empty(n + 1, np.uint32)
    col = np.empty(m.value, np.uint32)
    assert lib.hs_rmat_csr_host(scale, ef, a, b, c, seed, row.ctypes.data, col.ctypes.data,
                                C.byref(m)) == 0
    return row, col


def test_rmat_invariants_and_fixture(hs):
    fx = json.load(open(os.path.join(HERE, "golden", "rmat.json")))
    for key, exp in fx.items():
        scale, ef, seed = (int(x) for x in key.split(":"))
        row, col = _rmat(hs, scale, ef, seed)
        n = 1 << scale
        assert row[0] == 0 and row[-1] == col.size <= n * ef
        assert np.all(np.diff(row.astype(np.int64)) >= 0)
        for u in range(0, n, max(1, n // 64)):
            nb = col[row[u]:row[u + 1]]
            assert np.all(np.diff(nb.astype(np.int64)) > 0)  # sorted, no duplicates
            assert not np.any(nb == u)  # no self loops
        h = hashlib.sha256(row.tobytes() + col.tobytes()).hexdigest()
        assert (h, int(col.size)) == (exp["sha256"], exp["nedges"])
    # R-MAT skew: vertex 0 has the largest out-degree for a > b, c, d
    row, col = _rmat(hs, 12, 16, 1)
    deg = np.diff(row.astype(np.int64))
    assert deg.argmax() == 0


def test_gr_io_matches_oracle(hs, oracle, tmp_path):
    import ctypes as C
    lib = hs.load()
    row, col = _rmat(hs, 8, 8, 3)
    w = hs.gen_uniform_u32_host(col.size, 11, 1, 255)
    p = str(tmp_path / "g.gr")
    assert lib.hs_gr_write(p.encode(), row.size - 1, col.size, row.ctypes.data, col.ctypes.data,
                           w.ctypes.data) == 0
    r2, c2, w2 = oracle.gr_read(p)
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(w, w2)
    # and the product reader on the hand-packed fixture
    n, m, ety = C.c_uint64(), C.c_uint64(), C.c_uint64()
    tiny = os.path.join(HERE, "golden", "tiny.gr").encode()
    assert lib.hs_gr_read_header(tiny, C.byref(n), C.byref(m), C.byref(ety)) == 0
    assert (n.value, m.value, ety.value) == (3, 3, 4)
    r, c, ww = np.empty(4, np.uint32), np.empty(3, np.uint32), np.empty(3, np.uint32)
    assert lib.hs_gr_read(tiny, r.ctypes.data, c.ctypes.data, ww.ctypes.data) == 0
    assert r.tolist() == [0, 2, 3, 3] and c.tolist() == [1, 2, 2] and ww.tolist() == [5, 7, 1]
    # a corrupt destination id is rejected (LONESTAR/rt/src/csr_graph.cu:213)
    bad = bytearray(open(tiny, "rb").read())
    bad[32 + 24: 32 + 28] = (7).to_bytes(4, "little")
    pb = str(tmp_path / "bad.gr")
    open(pb, "wb").write(bad)
    assert lib.hs_gr_read(pb.encode(), r.ctypes.data, c.ctypes.data, ww.ctypes.data) != 0


def test_transpose(hs):
    lib = hs.load()
    row, col = _rmat(hs, 9, 8, 4)
    n = row.size - 1
    rrow = np.empty_like(row)
    rcol = np.empty_like(col)
    assert lib.hs_csr_transpose_host(n, col.size, row.ctypes.data, col.ctypes.data,
                                     rrow.ctypes.data, rcol.ctypes.data) == 0
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(row.astype(np.int64)))
    fwd = set(zip(src.tolist(), col.tolist()))
    dst = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rrow.astype(np.int64)))
    rev = set(zip(rcol.tolist(), dst.tolist()))
    assert fwd == rev


def test_python_front_end_tables(hs):
    assert hs.prim_id("atomicTreeBarrUniqCCG") == 45
    with pytest.raises(ValueError):
        hs.prim_id("bogus")
    assert hs.storage_locs(37, 100, 80, 32, 80) == 80 * 32 * 101


def test_serving_host_helpers(hs):
    """Host side of the serving loop without a device: the Poisson arrival
    times (sorted, deterministic per seed, mean gap 1/rate) and the latency
    summary of completed requests (percentiles of end-to-end and device
    service times)"""
    from hsync.serve import Server, poisson_arrivals
    a = poisson_arrivals(500.0, 20000, seed=5)
    assert np.all(np.diff(a) > 0) and a[0] > 0
    assert np.array_equal(a, poisson_arrivals(500.0, 20000, seed=5))
    assert not np.array_equal(a, poisson_arrivals(500.0, 20000, seed=6))
    gaps = np.diff(np.concatenate([[0.0], a]))
    assert abs(gaps.mean() * 500.0 - 1.0) < 0.03  # 20000 exponential gaps: std of the mean 0.7%
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05  # exponential: std = mean
    srv = Server.__new__(Server)  # no loop started: only the bookkeeping
    srv.lat = {}
    assert srv.percentiles() == {}
    srv.lat = {i: {"latency_s": (i + 1) * 1e-3, "service_s": (i + 1) * 5e-4, "status": 0} for i in range(100)}
    p = srv.percentiles()
    assert p["n"] == 100
    assert abs(p["p50_ms"] - 50.5) < 1e-9 and abs(p["service_p50_ms"] - 25.25) < 1e-9
    assert abs(p["p99_ms"] - 99.01) < 1e-9 and abs(p["service_p99_ms"] - 49.505) < 1e-9


def test_worklist_example_links_public_api_only(hs):
    """The worklist example application loads against libhs and exports its
    entry points; its build line names only the public include directory."""
    import ctypes as C
    lib = C.CDLL(os.path.join(os.path.dirname(hs.LIB_PATH), "examples", "libcc_labelprop.so"))
    assert hasattr(lib, "hs_example_cc") and hasattr(lib, "hs_example_cc_work_size")
    n = C.c_uint64(0)
    assert lib.hs_example_cc_work_size(C.c_uint32(1000), C.byref(n)) == 0 and n.value >= 3 * 4 * 1000
    src = open(os.path.join(os.path.dirname(hs.LIB_PATH), "examples", "cc_labelprop.hip")).read()
    assert '#include "hs_internal.hpp"' not in src
    mk = open(os.path.join(os.path.dirname(hs.LIB_PATH), "Makefile")).read()
    assert "EXFLAGS   := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -I../include" in mk


def test_serve_pagerank_tile_arithmetic(tmp_path):
    """The serving loop's PageRank tile placement and clamps
    (csrc/serve_tiles.hpp, the code its kernel runs) built for the host: on
    random graphs every tile start is placed once and the tiles' pieces cover
    every in-edge exactly once; with arbitrary tile starts every index stays in
    range.  A planted shift of the head pieces must be caught."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ absent")
    exe = str(tmp_path / "serve_tiles_check")
    r = subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(REPO, "tpds22-artifact_amd", "csrc"),
                        os.path.join(HERE, "serve_tiles_check.cpp"), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    ok = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert ok.returncode == 0 and "failures 0" in ok.stdout, ok.stdout + ok.stderr
    planted = subprocess.run([exe, "plant"], capture_output=True, text=True, timeout=300)
    assert planted.returncode == 1 and "failures 0" not in planted.stdout, planted.stdout
