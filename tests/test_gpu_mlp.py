"""The cost model's HIP kernels (forward, backward, Adam, inference, training)
against the float64 numpy restatement in tests/mlp_ref.py."""
import os

import numpy as np
import pytest

from tests import mlp_ref

pytestmark = pytest.mark.gpu


def setup(uhs, seed=3, n=300):
    m = uhs.MLP(seed)
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 2, (n, 20)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] + 0.1 * x[:, 3:].sum(1)).astype(np.float32)
    mean = x.mean(0); scale = x.std(0) + 0.5
    m.set_norm(mean, scale, float(t.mean()), float(t.std()))
    return m, x, t, mean, scale


def test_forward_and_gradient_match_numpy(gpu):
    m, x, t, mean, scale = setup(gpu)
    p = m.params().astype(np.float64)
    xs = (x.astype(np.float64) - mean) / scale
    ts = (t - t.mean()) / t.std()
    pred = m.predict(x)
    ref = mlp_ref.forward(p, xs)[-1][:, 0] * t.std() + t.mean()
    assert np.max(np.abs(pred - ref)) <= 1e-4 * max(1.0, np.abs(ref).max())
    g, loss = m.grad(x, t)
    rloss, rg = mlp_ref.loss_grad(p, xs, ts)
    assert abs(loss - rloss) <= 1e-5 * max(1.0, rloss)
    assert np.max(np.abs(g - rg)) <= 1e-4 * np.abs(rg).max()
    m.close()


def test_adam_step_matches_numpy(gpu):
    m, x, t, _, _ = setup(gpu, seed=5)
    p0 = m.params().astype(np.float64)
    g, _ = m.grad(x, t)
    m.adam_step(g, 1e-3, 1)
    p1, mo, v = mlp_ref.adam(p0, g.astype(np.float64), 0.0, 0.0, 1e-3, 1)
    assert np.max(np.abs(m.params() - p1)) <= 1e-6
    m.close()


def test_training_reduces_loss_and_save_load(gpu, tmp_path):
    m, x, t, _, _ = setup(gpu, seed=7, n=2048)
    _, l0 = m.grad(x, t)
    final = m.train(x, t, epochs=60, batch=128, lr=2e-3, fit_norm=True)
    _, l1 = m.grad(x, t)
    assert l1 < 0.2 * l0 and final < 0.2 * l0
    path = str(tmp_path / "m.bin")
    m.save(path)
    m2 = gpu.MLP(99)
    m2.load(path)
    assert np.array_equal(m2.predict(x[:50]), m.predict(x[:50]))
    m.close(); m2.close()
