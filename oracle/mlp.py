"""FP64 numpy restatement of the scheduler's cost model (TEST INFRASTRUCTURE ONLY).

The HIP library trains and evaluates a 3-layer perceptron (F -> 64 -> 64 -> 1,
ReLU) with its own kernels (tribit_amd/csrc/mlp.hip). This module is the
checker for those kernels: the same forward pass, loss, gradients and Adam
step in float64, plus a bfloat16 rounding model of the inference path.

No counterpart in the reference (the cost model is an addition). Pinned by:
  - central finite differences of the loss (tests/test_mlp_oracle.py);
  - a closed form: with ReLUs inactive the model is affine and the gradient
    of the bias b3 equals the mean residual;
  - Adam's first step moving every parameter by lr * sign(g) (Kingma & Ba).

Parameter layout (flat): W1[F][64] b1[64] W2[64][64] b2[64] W3[64] b3[1].
Loss: L = 1/(2B) * sum_b (y_b - t_b)^2.
"""
from __future__ import annotations

import numpy as np

H = 64


def shape(f: int) -> dict:
    o = {}
    off = 0
    for name, size in (("W1", f * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H), ("b3", 1)):
        o[name] = (off, size)
        off += size
    o["total"] = off
    return o


def unpack(p: np.ndarray, f: int) -> dict:
    s = shape(f)
    d = {}
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        a, n = s[name]
        d[name] = p[a:a + n]
    d["W1"] = d["W1"].reshape(f, H)
    d["W2"] = d["W2"].reshape(H, H)
    return d


def forward(p: np.ndarray, X: np.ndarray):
    f = X.shape[1]
    w = unpack(np.asarray(p, np.float64), f)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ w["W1"] + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"] + w["b2"], 0.0)
    y = h2 @ w["W3"] + w["b3"][0]
    return y, (X, h1, h2)


def loss(p: np.ndarray, X: np.ndarray, t: np.ndarray) -> float:
    y, _ = forward(p, X)
    return float(np.sum((y - np.asarray(t, np.float64)) ** 2) / (2 * X.shape[0]))


def grad(p: np.ndarray, X: np.ndarray, t: np.ndarray):
    """(loss, gradient) in float64."""
    f = X.shape[1]
    B = X.shape[0]
    w = unpack(np.asarray(p, np.float64), f)
    y, (X, h1, h2) = forward(p, X)
    dy = (y - np.asarray(t, np.float64)) / B
    g = {}
    g["W3"] = h2.T @ dy
    g["b3"] = np.array([dy.sum()])
    d2 = np.outer(dy, w["W3"]) * (h2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"].T) * (h1 > 0)
    g["W1"] = X.T @ d1
    g["b1"] = d1.sum(0)
    flat = np.concatenate([g[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
    return float(np.sum((y - t) ** 2) / (2 * B)), flat


def adam_step(p, g, m1, m2, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m1, m2)."""
    m1 = beta1 * m1 + (1 - beta1) * g
    m2 = beta2 * m2 + (1 - beta2) * g * g
    lr_t = lr * np.sqrt(1 - beta2 ** step) / (1 - beta1 ** step)
    return p - lr_t * m1 / (np.sqrt(m2) + eps), m1, m2


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def infer_bf16(p: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Inference with weights, inputs and hidden activations rounded to bf16 (float64 sums)."""
    f = X.shape[1]
    w = {k: v.astype(np.float64) for k, v in unpack(bf16(np.asarray(p, np.float32)), f).items()}
    x = bf16(np.asarray(X, np.float32)).astype(np.float64)
    h1 = bf16(np.maximum(x @ w["W1"] + w["b1"], 0.0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ w["W2"] + w["b2"], 0.0).astype(np.float32)).astype(np.float64)
    return h2 @ w["W3"] + w["b3"][0]
