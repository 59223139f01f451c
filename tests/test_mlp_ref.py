"""Pins the numpy MLP restatement (the checker of the GPU kernels) by
central finite differences, and the feature encoding of the C ABI."""
import numpy as np
import pytest

from tests import mlp_ref


def test_numpy_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = sum(i * o + o for i, o in mlp_ref.SIZES)
    p = rng.normal(0, 0.2, n)
    xs = rng.normal(0, 1, (9, 20))
    ts = rng.normal(0, 1, 9)
    loss, g = mlp_ref.loss_grad(p, xs, ts)
    for k in rng.choice(n, 40, replace=False):
        e = np.zeros(n); e[k] = 1e-6
        num = (mlp_ref.loss_grad(p + e, xs, ts)[0] - mlp_ref.loss_grad(p - e, xs, ts)[0]) / 2e-6
        assert abs(num - g[k]) <= 1e-6 * max(1.0, abs(num)), k


def test_feature_encoding(uhs):
    L = np.log2
    # 3D radius-2 box (125 terms), two planes of prefetch, non-temporal stores, depth 4
    f = uhs.mlp_features(3, 2, False, [64, 128, 256], (64, 8, 32, 2, 3), depth=4)
    assert np.allclose(f, [1, 0, 1, 0, 0, 6, 7, 8, 6, 3, 5, 8, 2, 0, 0, 1, L(125), 0, 1, 1])
    # 2D 5-point star
    f = uhs.mlp_features(2, 1, True, [1024, 2048], (512, 1, 64, 1, 0))
    assert np.allclose(f, [0, 1, 0, 0, 1, 10, 11, 0, 9, 0, 6, 9, 1, 1, 0, 0, L(5), 0, 0, 0])
    # two steps per launch: depth 2, threads over the ghost-zone tile (64/2 * 8)
    f = uhs.mlp_features(3, 2, True, [64, 128, 256], (60, 4, 64, 2, 33))
    assert np.allclose(f, [1, 0, 1, 0, 1, 6, 7, 8, L(60), 2, 6, 8, 2, 0, 1, 0, L(13), 0, 1, 0])
    # two rows per thread + two planes in flight: 64 * 16 / 2 threads, variant 6;
    # FP32 compensated
    f = uhs.mlp_features(3, 1, False, [64, 128, 256], (64, 16, 128, 1, 80), dtype=2)
    assert np.allclose(f, [1, 1, 0, 0, 0, 6, 7, 8, 6, 4, 7, 9, 1, 1, 0, 0, L(27), 2, 0, 6])


def test_parameter_count(uhs):
    assert uhs.lib().uhs_mlp_num_params() == sum(i * o + o for i, o in mlp_ref.SIZES) == 35841
