"""Float64 restatement of the plan-choice MLP (12 -> 64 -> 64 -> 3, ReLU).

TEST INFRASTRUCTURE ONLY: the checker of the library's FP32 forward,
backward and Adam kernels.  The cost model has no counterpart in the
reference (which decides plans by fixed rules, ``src/column.cpp:6102-6111``
and ``src/index.cpp:6640-6783``), so this restatement is pinned by what the
mathematics fixes instead: central finite differences of the loss, closed
forms for a network whose weights are zero, and Adam's first step, which
moves every parameter by lr * g / (|g| + eps).

Parameter layout (flat): W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3].
Loss: L = sum_i sum_o w_io e_io^2 / sum |mask|, e = score - y; mask > 0 is an
exact label of weight mask, mask < 0 a lower bound of weight -mask (e counts
only when negative), mask == 0 no label.
"""
from __future__ import annotations

import numpy as np

NF, H1, H2, NO = 12, 64, 64, 3
SHAPES = [("W1", (H1, NF)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)), ("W3", (NO, H2)),
          ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(params):
    p = np.asarray(params, np.float64)
    assert p.size == NPARAM
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def normalise(x, mu, sd):
    sd = np.where(np.asarray(sd, np.float64) > 0, sd, 1.0)
    return (np.asarray(x, np.float64) - mu) / sd


def forward(params, xn):
    """Scores and the hidden activations of normalised inputs xn[n, 12]."""
    P = unpack(params)
    a1 = np.maximum(xn @ P["W1"].T + P["b1"], 0.0)
    a2 = np.maximum(a1 @ P["W2"].T + P["b2"], 0.0)
    out = a2 @ P["W3"].T + P["b3"]
    return out, a1, a2


def residual(out, y, mask):
    e = out - y
    e = np.where((mask < 0) & (e > 0), 0.0, e)
    return e, np.abs(mask)


def loss(params, xn, y, mask):
    out, _, _ = forward(params, xn)
    e, w = residual(out, y, mask)
    m = w.sum()
    return float((w * e * e).sum() / m) if m > 0 else 0.0


def grad(params, xn, y, mask):
    """(loss, dL/dparams) by backpropagation."""
    P = unpack(params)
    out, a1, a2 = forward(params, xn)
    e, w = residual(out, y, mask)
    m = w.sum()
    inv = 1.0 / m if m > 0 else 0.0
    d3 = 2.0 * w * e * inv
    d2 = (d3 @ P["W3"]) * (a2 > 0)
    d1 = (d2 @ P["W2"]) * (a1 > 0)
    g = {"W3": d3.T @ a2, "b3": d3.sum(0), "W2": d2.T @ a1, "b2": d2.sum(0), "W1": d1.T @ xn,
         "b1": d1.sum(0)}
    return float((w * e * e).sum() * inv), pack(g)


def adam_step(params, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015, Algorithm 1) step t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
