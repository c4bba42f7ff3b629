"""The cost model's HIP kernels (tribit_amd/csrc/mlp.hip) against the FP64
oracle (oracle/mlp.py): loss and gradients of the FP32 training kernels, an
Adam step, training that lowers the loss, and BF16 inference."""
import numpy as np
import pytest

from oracle import mlp as ref

pytestmark = pytest.mark.gpu


def data(f, b, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (b, f)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] + 0.1 * rng.normal(size=b)).astype(np.float32)
    return X, t


@pytest.mark.parametrize("f,b", [(16, 1), (16, 257), (5, 1024)])
def test_gradients_match_fp64_oracle(gpu_ctx, f, b):
    X, t = data(f, b, f + b)
    with gpu_ctx.mlp(f, seed=7) as m:
        p = m.params().astype(np.float64)
        loss, g = m.grad(X, t)
        rl, rg = ref.grad(p, X.astype(np.float64), t.astype(np.float64))
        assert abs(loss - rl) <= 1e-5 * max(1.0, abs(rl))
        # FP32 sums of <= 1024 terms: relative error well under 1e-4 of the gradient's scale
        assert np.max(np.abs(g - rg)) <= 1e-4 * max(1.0, np.max(np.abs(rg)))


def test_adam_step_matches_oracle(gpu_ctx):
    f, b = 8, 64
    X, t = data(f, b, 3)
    with gpu_ctx.mlp(f, seed=2) as m:
        p0 = m.params().astype(np.float64)
        m.train(X, t, epochs=1, batch=b, lr=1e-2)  # one batch == one Adam step
        p1 = m.params().astype(np.float64)
        _, g = ref.grad(p0, X.astype(np.float64), t.astype(np.float64))
        want, _, _ = ref.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
        assert np.max(np.abs(p1 - want)) <= 2e-5


def test_training_lowers_loss_and_is_deterministic(gpu_ctx):
    f = 16
    X, t = data(f, 4096, 11)
    with gpu_ctx.mlp(f, seed=5) as m, gpu_ctx.mlp(f, seed=5) as m2:
        before = ref.loss(m.params().astype(np.float64), X, t)
        after = m.train(X, t, epochs=30, batch=256, lr=3e-3, seed=9)
        m2.train(X, t, epochs=30, batch=256, lr=3e-3, seed=9)
        assert np.array_equal(m.params(), m2.params())
        final = ref.loss(m.params().astype(np.float64), X, t)
        assert final < 0.25 * before and abs(final - after) < 0.5 * after + 1e-3


def test_bf16_inference(gpu_ctx):
    f = 16
    X, t = data(f, 3000, 12)
    with gpu_ctx.mlp(f, seed=4) as m:
        m.train(X, t, epochs=5, batch=256, lr=3e-3)
        p = m.params()
        y = m.infer(X)
        want = ref.infer_bf16(p, X)
        fp64, _ = ref.forward(p.astype(np.float64), X.astype(np.float64))
        # same bf16 rounding points; FP32 vs FP64 accumulation only
        assert np.max(np.abs(y - want)) <= 5e-3 * max(1.0, np.max(np.abs(want)))
        # and bf16 stays close to the full-precision model
        assert np.max(np.abs(y - fp64)) <= 0.05 * max(1.0, np.max(np.abs(fp64)))
