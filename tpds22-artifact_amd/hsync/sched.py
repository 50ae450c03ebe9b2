"""Learned push/pull schedule for the multi-source BFS.

Every BFS level is a sample: its features (frontier vertices nf, frontier
out-edges mf, open in-edges mu, discoveries of the previous level, graph size,
level) are schedule-invariant -- the frontier at level L is the same whatever
directions earlier levels took -- so one all-push and one all-pull run of a
graph measure, for every level, the time of both choices on the same state.
The label is the faster direction; the sample weight is the time the wrong
choice would cost.  The MLP (8 -> 64 -> 64 -> 64 -> 2) is trained with libhs's
HIP forward/backward kernels (Adam) and evaluated inside the BFS kernel; the
plain heuristic stays the fallback.  Any schedule gives the same levels.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib as L
from . import graph as G
from .api import _ptr, _stream

N_FEATURES = 8
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "sched_weights.json")


def _torch():
    import torch
    return torch


def features(nf, mf, mu, disc, n, m, level) -> np.ndarray:
    """the same features as ms_features() in csrc/graph.hip (float32)"""
    lg = lambda x: np.log2(np.float32(1.0) + np.float32(x))  # noqa: E731
    return np.array([lg(nf) / 24, lg(mf) / 32, lg(mu) / 32, lg(disc) / 32, lg(n) / 32, lg(m) / 32,
                     np.float32(level) / 32, (lg(mf) - lg(mu)) / 8], dtype=np.float32)


def collect(arena, g, srcs, barrier="xcd", repeats=2):
    """per-level samples of one graph and source batch: features, push time, pull time"""
    best = {}
    for sched in ("push", "pull"):
        per_level = None
        for _ in range(repeats):
            tr = G.new_trace()
            G.msbfs(arena, g, srcs, schedule=sched, barrier=barrier, trace=tr)
            recs = G.decode_trace(tr)
            if per_level is None:
                per_level = recs
            else:  # keep the faster repeat per level (less noise)
                for a, b in zip(per_level, recs):
                    a["seconds"] = min(a["seconds"], b["seconds"])
        best[sched] = per_level
    samples = []
    for L_, (a, b) in enumerate(zip(best["push"], best["pull"]), start=1):
        assert (a["nf"], a["mf"], a["mu"], a["disc"]) == (b["nf"], b["mf"], b["mu"], b["disc"])
        samples.append({"x": features(a["nf"], a["mf"], a["mu"], a["disc"], g.n, g.m, L_).tolist(),
                        "t_push": a["seconds"], "t_pull": b["seconds"], "level": L_,
                        "nf": a["nf"], "mf": a["mf"], "mu": a["mu"]})
    return samples


def to_arrays(samples):
    X = np.array([s["x"] for s in samples], dtype=np.float32)
    tp = np.array([s["t_push"] for s in samples])
    tl = np.array([s["t_pull"] for s in samples])
    y = (tl < tp).astype(np.int32)
    w = np.abs(tp - tl).astype(np.float32) + np.float32(1e-7)
    return X, y, w


class MLP:
    """parameters on the device + the libhs training / inference kernels"""

    def __init__(self, seed: int = 0, params=None):
        torch = _torch()
        lib = L.load()
        self.nparams = lib.hs_mlp_param_count()
        if params is not None:
            self.w = torch.as_tensor(np.asarray(params, np.float32)).cuda()
            assert self.w.numel() == self.nparams
        else:
            self.w = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
            L.check(lib.hs_mlp_init(_ptr(self.w), seed, _stream()), "hs_mlp_init")
        self.scratch = torch.zeros(lib.hs_mlp_scratch_floats(), dtype=torch.float32, device="cuda")
        self.step = 0

    def logits(self, X):
        torch = _torch()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        out = torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda")
        L.check(L.load().hs_mlp_forward(_ptr(self.w), _ptr(x), x.shape[0], _ptr(out), _stream()),
                "hs_mlp_forward")
        return out

    def grad(self, X, y, w):
        torch = _torch()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        yy = torch.as_tensor(y, dtype=torch.int32).cuda()
        ww = torch.as_tensor(w, dtype=torch.float32).cuda()
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        L.check(L.load().hs_mlp_grad(_ptr(self.w), _ptr(x), _ptr(yy), _ptr(ww), x.shape[0],
                                     float(ww.sum().item()), _ptr(g), _ptr(loss), _ptr(self.scratch),
                                     _stream()), "hs_mlp_grad")
        return g, float(loss.item())

    def fit(self, X, y, w, epochs=400, lr=3e-3):
        torch = _torch()
        lib = L.load()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        yy = torch.as_tensor(y, dtype=torch.int32).cuda()
        ww = torch.as_tensor(w, dtype=torch.float32).cuda()
        wsum = float(ww.sum().item())
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        hist = []
        for _ in range(epochs):
            self.step += 1
            L.check(lib.hs_mlp_train_step(_ptr(self.w), _ptr(x), _ptr(yy), _ptr(ww), x.shape[0], wsum,
                                          lr, self.step, _ptr(loss), _ptr(self.scratch), _stream()),
                    "hs_mlp_train_step")
            if self.step % 50 == 0:
                hist.append(float(loss.item()))
        return hist

    def predict(self, X):
        return self.logits(X).argmax(1).cpu().numpy().astype(np.int32)

    def save(self, path=WEIGHTS_PATH, meta=None):
        json.dump({"params": self.w.cpu().numpy().tolist(), "meta": meta or {}}, open(path, "w"))

    @staticmethod
    def load(path=WEIGHTS_PATH):
        d = json.load(open(path))
        return MLP(params=d["params"])


def heuristic_choice(nf, mf, mu):
    """the plain schedule of the kernel (heuristic_dir in csrc/graph.hip)"""
    return 1 if mf * 2 > mu else 0


def schedule_cost(samples, choices):
    return float(sum(s["t_pull"] if c else s["t_push"] for s, c in zip(samples, choices)))
