"""The command-line drop-ins (tpds22-artifact_amd/bin/) run with the reference
binaries' arguments and produce the reference's outputs, checked against the
oracle: allSyncPrims-1kernel (golden storage, "PASSED!"), reduce ("Sum" equals
"Correct sum"), scan ("Match: 1"), and bfs / sssp / pr on a .gr file (output
files vs the CPU oracle)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tpds22-artifact_amd", "bin")


def run(args, stdin=None, timeout=120):
    print("cli:", " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, args[0])] + list(args[1:]), input=stdin, capture_output=True,
                       text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("prim", ["atomicTreeBarrUniqSRB", "atomicTreeBarrUniqLocalExchSRB", "lfTreeBarrUniq",
                                  "atomicTreeBarrUniqCCG", "spinMutex", "faMutexUniq", "sleepMutex",
                                  "spinSemEBO10", "PRspinSem2", "spinSemUniq120"])
def test_allsyncprims_cli(prim):
    rc, out = run(["allSyncPrims-1kernel", prim, "10", "24", "2"])
    assert rc == 0 and "PASSED!" in out, out[-2000:]


def test_reduce_cli():
    for kind in ("xcd", "flat", "coop"):
        rc, out = run(["reduce", str((1 << 20) + 3), "256", kind])
        assert rc == 0, out[-2000:]
        sums = [int(line.split(":")[1]) for line in out.splitlines() if line.startswith(("Sum:", "Correct sum:"))]
        assert len(sums) == 2 and sums[0] == sums[1], out[-2000:]


def test_reduce_cli_graph_replay():
    # the launch captured once (hs_graph_capture_*) and replayed 50 times
    rc, out = run(["reduce", str(1 << 22), "256", "xcd", "50"])
    assert rc == 0, out[-2000:]
    sums = [int(line.split(":")[1]) for line in out.splitlines() if line.startswith(("Sum:", "Correct sum:"))]
    assert sums == [1 << 22, 1 << 22], out[-2000:]


def test_scan_cli():
    for args in (["scan", "1000003"], ["scan", "65536", "1"]):
        rc, out = run(args)
        assert rc == 0 and "Match: 1" in out, out[-2000:]


@pytest.fixture(scope="module")
def grfile(tmp_path_factory):
    import hsync.graph as G
    row, col = G.rmat_csr(12, 8, 41)
    w = G.weights(col.size, 42)
    path = str(tmp_path_factory.mktemp("gr") / "rmat12.gr")
    G.write_gr(path, row, col, w)
    return path, row, col, w


def _read_pairs(path):
    out = {}
    for line in open(path):
        a, b = line.split()
        out[int(a)] = 0x7FFFFFFF if b == "INF" else int(b)
    return np.array([out[i] for i in range(len(out))], dtype=np.int64)


def test_bfs_sssp_cli(grfile, oracle, tmp_path):
    path, row, col, w = grfile
    out = str(tmp_path / "bfs.txt")
    rc, log = run(["bfs", "-s", "3", "-o", out, path], stdin="2\n")
    assert rc == 0, log[-2000:]
    assert np.array_equal(_read_pairs(out), oracle.bfs(row, col, 3).astype(np.int64))
    out = str(tmp_path / "bfs_lb.txt")  # -l: the load-balanced expansion
    rc, log = run(["bfs", "-l", "-s", "3", "-o", out, path], stdin="1\n")
    assert rc == 0, log[-2000:]
    assert np.array_equal(_read_pairs(out), oracle.bfs(row, col, 3).astype(np.int64))
    out = str(tmp_path / "sssp.txt")
    rc, log = run(["sssp", "-d", "100", "-s", "5", "-o", out, path], stdin="2\n")
    assert rc == 0, log[-2000:]
    assert np.array_equal(_read_pairs(out), oracle.sssp(row, col, w, 5).astype(np.int64))


def test_pr_cli(grfile, oracle, tmp_path):
    path, row, col, w = grfile
    out = str(tmp_path / "pr.txt")
    rc, log = run(["pr", "-x", "100", "-o", out, path], stdin="2\n")
    assert rc == 0, log[-2000:]
    ref, _ = oracle.pagerank(row, col, alpha=0.85, tol=1e-14)
    got = np.zeros(row.size - 1)
    for line in open(out):
        v, r = line.split()
        got[int(v)] = float(r)
    assert np.abs(got - ref).max() < 1e-6
