"""Learned schedule for the multi-source BFS.

Every BFS level is a sample.  Its features (frontier vertices nf, frontier
out-edges mf, open in-edges mu, discoveries of the previous level, graph size,
level) are schedule-invariant -- the frontier at level L is the same whatever
earlier levels did -- so one run per action with that action forced at every
level measures, for every level, the time of each of the 8 actions on the same
state.  An action is the level's direction (push / pull), whether only half of
the workgroups expand it, and the barrier poll backoff after it
(action = dir | half << 1 | long_backoff << 2).

The MLP (8 -> 64 -> 64 -> 64 -> 8) is trained with libhs's HIP forward/backward
kernels (Adam) on the expected time of its softmax policy over those measured
times, and evaluated inside the BFS kernel.  The plain heuristic stays the
judge and the fallback: the network's action replaces the heuristic's only when
its score leads by a margin, chosen per graph-size band on validation graphs of
that band so that the learned schedule is no slower there than the heuristic; a
margin of +inf is the plain schedule.  Any schedule gives the same levels.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import _lib as L
from . import graph as G
from .api import _ptr, _stream

N_FEATURES = 8
N_ACTIONS = 8
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "sched_weights.json")


def _torch():
    import torch
    return torch


def action(direction: int, half: bool = False, long_backoff: bool = False) -> int:
    return int(direction) | (int(half) << 1) | (int(long_backoff) << 2)


def features(nf, mf, mu, disc, n, m, level) -> np.ndarray:
    """the same features as ms_features() in csrc/graph.hip (float32): the level's
    counts relative to the graph, mean degree, level, edge balance, graph size"""
    lg = lambda x: np.log2(np.float32(1.0) + np.float32(x))  # noqa: E731
    ln, lm = lg(n), lg(m)
    return np.array([(lg(nf) - ln) / 32, (lg(mf) - lm) / 32, (lg(mu) - lm) / 32, (lg(disc) - ln) / 32,
                     (lm - ln) / 8, np.float32(level) / 32, (lg(mf) - lg(mu)) / 8, ln / 32], dtype=np.float32)


def heuristic_choice(nf, mf, mu):
    """the plain schedule's action (heuristic_dir in csrc/graph.hip): pull when the
    frontier's out-edges exceed half of the unvisited vertices' in-edges; all
    workgroups, short backoff"""
    return 1 if mf * 2 > mu else 0


def collect(arena, g, srcs, barrier="xcd", repeats=2, actions=range(N_ACTIONS)):
    """per-level samples of one graph and source batch: features and the time of
    every action (seconds; inf for actions not measured)"""
    per_action = {}
    for a in actions:
        best = None
        for _ in range(repeats):
            tr = G.new_trace()
            G.msbfs(arena, g, srcs, schedule="forced", forced=[a] * G.MAX_LEVELS, barrier=barrier, trace=tr)
            recs = G.decode_trace(tr)
            if best is None:
                best = recs
            else:  # keep the faster repeat per level (less noise)
                for x, y in zip(best, recs):
                    x["seconds"] = min(x["seconds"], y["seconds"])
        per_action[a] = best
    ref = per_action[next(iter(per_action))]
    samples = []
    for L_, r in enumerate(ref, start=1):
        t = np.full(N_ACTIONS, np.inf)
        for a, recs in per_action.items():
            q = recs[L_ - 1]
            assert (q["nf"], q["mf"], q["mu"], q["disc"]) == (r["nf"], r["mf"], r["mu"], r["disc"])
            t[a] = q["seconds"]
        samples.append({"x": features(r["nf"], r["mf"], r["mu"], r["disc"], g.n, g.m, L_).tolist(),
                        "t": t.tolist(), "level": L_, "nf": r["nf"], "mf": r["mf"], "mu": r["mu"],
                        "t_push": float(t[0]), "t_pull": float(t[1])})
    return samples


def to_arrays(samples):
    """features, normalised costs [n][8] (unmeasured actions cost the level's worst
    time twice) and the normalisation (the sum of the per-level best times)"""
    X = np.array([s["x"] for s in samples], dtype=np.float32)
    T = np.array([s["t"] for s in samples], dtype=np.float64)
    worst = np.where(np.isfinite(T), T, -np.inf).max(1, keepdims=True)
    T = np.where(np.isfinite(T), T, 2 * worst)
    scale = T.min(1).sum()
    return X, (T / scale).astype(np.float32), float(scale)


def plain_actions(samples):
    return [heuristic_choice(s["nf"], s["mf"], s["mu"]) for s in samples]


def schedule_cost(samples, choices):
    return float(sum(s["t"][c] for s, c in zip(samples, choices)))


class MLP:
    """parameters on the device, followed by the judge (margin for small graphs,
    margin for large graphs, the size threshold in log2(n + 1)), and the libhs
    training / inference kernels"""

    def __init__(self, seed: int = 0, params=None, margin: float = 0.0):
        torch = _torch()
        lib = L.load()
        self.nparams = lib.hs_mlp_param_count()
        self.nactions = lib.hs_mlp_actions()
        self.w = torch.zeros(self.nparams + 3, dtype=torch.float32, device="cuda")
        if params is not None:
            p = np.asarray(params, np.float32)
            assert p.size == self.nparams
            self.w[:self.nparams] = torch.as_tensor(p).cuda()
        else:
            L.check(lib.hs_mlp_init(_ptr(self.w), seed, _stream()), "hs_mlp_init")
        self.set_margin(margin)
        self.scratch = torch.zeros(lib.hs_mlp_scratch_floats(), dtype=torch.float32, device="cuda")
        self.step = 0

    @property
    def params(self):
        return self.w[:self.nparams]

    def set_judge(self, margin_small: float, margin_large: float, threshold: float) -> None:
        """margin_small for graphs with log2(n + 1) <= threshold, margin_large above"""
        self.judge = (float(margin_small), float(margin_large), float(threshold))
        self.margin = float(margin_small)
        self.w[self.nparams:self.nparams + 3] = _torch().tensor(self.judge, dtype=_torch().float32)

    def set_margin(self, margin: float) -> None:
        """one margin for every graph size"""
        self.set_judge(margin, margin, math.inf)

    def _margins(self, samples):
        small, large, thr = self.judge
        lgn = np.array([s["x"][7] for s in samples], dtype=np.float32) * np.float32(32.0)
        return np.where(lgn <= np.float32(thr), small, large)

    def logits(self, X):
        torch = _torch()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        out = torch.empty((x.shape[0], self.nactions), dtype=torch.float32, device="cuda")
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

    def grad_cost(self, X, cost, norm=1.0):
        torch = _torch()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        c = torch.as_tensor(cost, dtype=torch.float32).cuda().contiguous()
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        L.check(L.load().hs_mlp_grad_cost(_ptr(self.w), _ptr(x), _ptr(c), x.shape[0], norm, _ptr(g),
                                          _ptr(loss), _ptr(self.scratch), _stream()), "hs_mlp_grad_cost")
        return g, float(loss.item())

    def fit(self, X, y, w, epochs=400, lr=3e-3):
        """cross-entropy against one action label per sample"""
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

    def fit_cost(self, X, cost, epochs=800, lr=3e-3):
        """the expected normalised time of the softmax policy (1.0 = the per-level best)"""
        torch = _torch()
        lib = L.load()
        x = torch.as_tensor(X, dtype=torch.float32).cuda().contiguous()
        c = torch.as_tensor(cost, dtype=torch.float32).cuda().contiguous()
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        hist = []
        for _ in range(epochs):
            self.step += 1
            L.check(lib.hs_mlp_train_step_cost(_ptr(self.w), _ptr(x), _ptr(c), x.shape[0], 1.0, lr, self.step,
                                               _ptr(loss), _ptr(self.scratch), _stream()),
                    "hs_mlp_train_step_cost")
            if self.step % 50 == 0:
                hist.append(float(loss.item()))
        return hist

    def predict(self, X):
        return self.logits(X).argmax(1).cpu().numpy().astype(np.int32)

    def judged(self, samples, margin=None):
        """the actions the kernel takes: the network's best when it leads the
        heuristic's action by at least the margin (the judge's margin of each
        sample's size band, or the one given), else the heuristic's"""
        margins = self._margins(samples) if margin is None else np.full(len(samples), margin)
        X = np.array([s["x"] for s in samples], dtype=np.float32)
        o = self.logits(X).cpu().numpy().astype(np.float64)
        plain = plain_actions(samples)
        best = o.argmax(1)
        lead = o[np.arange(len(samples)), best] - o[np.arange(len(samples)), plain]
        return [int(b) if ld >= m else int(p) for b, ld, p, m in zip(best, lead, plain, margins)]

    def choose_margin(self, val_graphs, candidates=(0.0, 0.25, 0.5, 1.0, 2.0, 4.0)):
        """the margin whose judged schedule is fastest on validation graphs (a list
        of per-graph sample lists, as collect returns them): the mean over the
        graphs of the judged schedule's level time over the plain schedule's, so
        every graph counts alike whatever its size; +inf (the plain schedule)
        unless some margin's mean is below 1.  Returns (margin, mean ratio)."""
        if val_graphs and isinstance(val_graphs[0], dict):
            val_graphs = [val_graphs]
        plain = [schedule_cost(g, plain_actions(g)) for g in val_graphs]

        def ratio(m):
            return float(np.mean([schedule_cost(g, self.judged(g, m)) / p for g, p in zip(val_graphs, plain)]))

        best_m, best_r = math.inf, 1.0
        for m in candidates:
            r = ratio(m)
            if r < best_r:
                best_m, best_r = m, r
        self.set_margin(best_m)
        return best_m, best_r

    def choose_judge(self, val_small, val_large, threshold, candidates=(0.0, 0.25, 0.5, 1.0, 2.0, 4.0)):
        """a margin per graph-size band, each chosen (as choose_margin) on the
        validation graphs of its band: val_small below log2(n + 1) = threshold,
        val_large above.  Returns (margin_small, ratio_small, margin_large, ratio_large)."""
        ms, rs = self.choose_margin(val_small, candidates)
        ml, rl = self.choose_margin(val_large, candidates)
        self.set_judge(ms, ml, threshold)
        return ms, rs, ml, rl

    def save(self, path=WEIGHTS_PATH, meta=None):
        json.dump({"params": self.params.cpu().numpy().tolist(), "judge": list(self.judge), "meta": meta or {}},
                  open(path, "w"))

    @staticmethod
    def load(path=WEIGHTS_PATH):
        d = json.load(open(path))
        m = MLP(params=d["params"])
        m.set_judge(*d.get("judge", (0.0, 0.0, math.inf)))
        return m
