"""Pins the FP64 cost-model oracle (oracle/mlp.py) before it checks the HIP
training kernels: finite differences, an affine closed form, Adam's first step
and the bfloat16 rounding rule."""
import numpy as np

from oracle import mlp


def make(f=12, b=64, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.shape(f)["total"])
    X = rng.normal(0, 1, (b, f))
    t = rng.normal(0, 1, b)
    return p, X, t


def test_gradient_matches_finite_differences():
    p, X, t = make()
    _, g = mlp.grad(p, X, t)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (mlp.loss(p + e, X, t) - mlp.loss(p - e, X, t)) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), i


def test_affine_closed_form():
    """With every hidden pre-activation positive the model is affine: dL/db3 = mean residual."""
    f = 6
    s = mlp.shape(f)
    p = np.zeros(s["total"])
    a, n = s["b1"]
    p[a:a + n] = 100.0  # ReLUs stay active for small inputs
    a, n = s["b2"]
    p[a:a + n] = 100.0
    rng = np.random.default_rng(2)
    a, n = s["W2"]
    p[a:a + n] = rng.normal(0, 0.01, n)
    a, n = s["W3"]
    p[a:a + n] = rng.normal(0, 0.1, n)
    X = rng.normal(0, 0.1, (32, f))
    t = rng.normal(0, 1, 32)
    y, _ = mlp.forward(p, X)
    _, g = mlp.grad(p, X, t)
    assert np.isclose(g[s["b3"][0]], np.mean(y - t))
    # dL/dW3 = h2^T r / B with h2 = X W1 + b1 ... affine in the inputs
    _, (x, h1, h2) = mlp.forward(p, X)
    a, n = s["W3"]
    assert np.allclose(g[a:a + n], h2.T @ (y - t) / 32)


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(3)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    q, m1, m2 = mlp.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # exact up to eps / (sqrt(1 - beta2) |g|)
    assert np.allclose(p - q, 1e-3 * np.sign(g), rtol=1e-6 / np.abs(g).min() + 1e-9)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.140625, 0.0], np.float32)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 3.140625, 0.0], np.float32)
    assert np.array_equal(mlp.bf16(x), want)
    y = np.random.default_rng(4).normal(size=1000).astype(np.float32)
    r = mlp.bf16(y)
    assert np.all(np.abs(r - y) <= np.abs(y) * 2 ** -8)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
