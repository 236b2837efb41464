"""FP64 CPU restatement of the learned-refinement network.

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels in
mepinterp_amd/csrc/mep_mlp.hip (forward pass, loss, gradients).  The
reference has no learned component, so there is nothing of it to follow: the
network is defined by the C-ABI documentation (include/mepinterp.h) and
restated here in plain numpy.  Its analytic gradient is pinned by central
finite differences of its own loss (tests/test_oracle_mlp.py).

Parameter layout (flat, FP32 on the GPU, FP64 here):
  W1[FP][H] b1[H] W2[H][H] b2[H] W3[H][H] b3[H] W4[H][16] b4[16]
with FP = 6 n_freq rounded up to a multiple of 16 and column 0 of the last
layer the output.
"""
from __future__ import annotations

import numpy as np

OUTP = 16


def dims(n_freq: int, hidden: int):
    fp = ((6 * n_freq + 15) // 16) * 16
    sizes = [("W1", (fp, hidden)), ("b1", (hidden,)), ("W2", (hidden, hidden)), ("b2", (hidden,)),
             ("W3", (hidden, hidden)), ("b3", (hidden,)), ("W4", (hidden, OUTP)), ("b4", (OUTP,))]
    return fp, sizes


def n_params(n_freq: int, hidden: int) -> int:
    return sum(int(np.prod(s)) for _, s in dims(n_freq, hidden)[1])


def unpack(params, n_freq: int, hidden: int):
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shp in dims(n_freq, hidden)[1]:
        sz = int(np.prod(shp))
        out[name] = p[o:o + sz].reshape(shp)
        o += sz
    assert o == p.size
    return out


def features(kpts, n_freq: int):
    """x[r, a*2nf + 2m + {0,1}] = sin / cos(2 pi (m+1) k_a), zero-padded."""
    k = np.asarray(kpts, dtype=np.float64).reshape(3, -1)
    fp, _ = dims(n_freq, 16)
    x = np.zeros((k.shape[1], fp))
    for a in range(3):
        for m in range(n_freq):
            ang = 2.0 * np.pi * (m + 1) * k[a]
            x[:, a * 2 * n_freq + 2 * m] = np.sin(ang)
            x[:, a * 2 * n_freq + 2 * m + 1] = np.cos(ang)
    return x


def forward(params, kpts, n_freq: int, hidden: int, keep=False):
    w = unpack(params, n_freq, hidden)
    x = features(kpts, n_freq)
    h1 = np.tanh(x @ w["W1"] + w["b1"])
    h2 = np.tanh(h1 @ w["W2"] + w["b2"])
    h3 = np.tanh(h2 @ w["W3"] + w["b3"])
    y = h3 @ w["W4"] + w["b4"]
    if keep:
        return y[:, 0], (x, h1, h2, h3, w)
    return y[:, 0]


def loss(params, kpts, target, n_freq: int, hidden: int) -> float:
    y = forward(params, kpts, n_freq, hidden)
    e = y - np.asarray(target, dtype=np.float64)
    return float(np.mean(e * e))


def grad(params, kpts, target, n_freq: int, hidden: int):
    """Analytic gradient of `loss` (backpropagation), flat, FP64."""
    y, (x, h1, h2, h3, w) = forward(params, kpts, n_freq, hidden, keep=True)
    n = x.shape[0]
    dy = np.zeros((n, OUTP))
    dy[:, 0] = 2.0 * (y - np.asarray(target, dtype=np.float64)) / n
    g = {"W4": h3.T @ dy, "b4": dy.sum(axis=0)}
    dz3 = (dy @ w["W4"].T) * (1.0 - h3 * h3)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(axis=0)
    dz2 = (dz3 @ w["W3"].T) * (1.0 - h2 * h2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(axis=0)
    dz1 = (dz2 @ w["W2"].T) * (1.0 - h1 * h1)
    g["W1"], g["b1"] = x.T @ dz1, dz1.sum(axis=0)
    return np.concatenate([g[name].ravel() for name, _ in dims(n_freq, hidden)[1]])


def grad_fd(params, kpts, target, n_freq: int, hidden: int, index, h=1e-6):
    """Central finite differences of `loss` for the parameters in `index`."""
    p = np.asarray(params, dtype=np.float64).copy()
    out = np.zeros(len(index))
    for j, i in enumerate(index):
        s = h * max(1.0, abs(p[i]))
        p0 = p[i]
        p[i] = p0 + s
        lp = loss(p, kpts, target, n_freq, hidden)
        p[i] = p0 - s
        lm = loss(p, kpts, target, n_freq, hidden)
        p[i] = p0
        out[j] = (lp - lm) / (2.0 * s)
    return out
