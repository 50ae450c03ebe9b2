"""FP64 numpy restatement of the schedule MLP (8 -> 64 -> 64 -> 64 -> 8, ReLU),
its sample-weighted softmax cross-entropy and its expected-cost loss -- the
checker for the HIP training kernels.  Test infrastructure only."""
import numpy as np

IN, H, OUT = 8, 64, 8


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0

    def take(*shape):
        nonlocal o
        k = int(np.prod(shape))
        a = p[o:o + k].reshape(shape)
        o += k
        return a

    W1, b1 = take(H, IN), take(H)
    W2, b2 = take(H, H), take(H)
    W3, b3 = take(H, H), take(H)
    W4, b4 = take(OUT, H), take(OUT)
    assert o == p.size
    return W1, b1, W2, b2, W3, b3, W4, b4


def forward(p, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    z3 = h2 @ W3.T + b3
    h3 = np.maximum(z3, 0)
    o = h3 @ W4.T + b4
    return o, (X, z1, h1, z2, h2, z3, h3)


def _softmax(o):
    m = o.max(1, keepdims=True)
    e = np.exp(o - m)
    return e / e.sum(1, keepdims=True)


def _backward(p, cache, d):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X, z1, h1, z2, h2, z3, h3 = cache
    gW4, gb4 = d.T @ h3, d.sum(0)
    dz3 = (d @ W4) * (z3 > 0)
    gW3, gb3 = dz3.T @ h2, dz3.sum(0)
    dz2 = (dz3 @ W3) * (z2 > 0)
    gW2, gb2 = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1, gb1 = dz1.T @ X, dz1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])


def loss_and_grad(p, X, y, w):
    """sum_s w_s * CE(softmax(o_s), y_s) / sum(w)"""
    o, cache = forward(p, X)
    w = np.asarray(w, np.float64) / np.sum(w)
    P = _softmax(o)
    y = np.asarray(y)
    loss = float(-(w * np.log(P[np.arange(len(y)), y])).sum())
    d = P.copy()
    d[np.arange(len(y)), y] -= 1.0
    d *= w[:, None]
    return loss, _backward(p, cache, d)


def cost_loss_and_grad(p, X, cost, norm):
    """sum_s sum_a softmax(o_s)_a * cost[s][a] / norm"""
    o, cache = forward(p, X)
    P = _softmax(o)
    cost = np.asarray(cost, np.float64)
    ec = (P * cost).sum(1, keepdims=True)
    loss = float(ec.sum() / norm)
    d = P * (cost - ec) / norm
    return loss, _backward(p, cache, d)


def nparams():
    return H * IN + H + H * H + H + H * H + H + OUT * H + OUT
