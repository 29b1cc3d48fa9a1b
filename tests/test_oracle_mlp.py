"""Pins the float64 MLP oracle (CPU only): analytic gradient against central
finite differences, Adam against its closed form on a quadratic, BF16 rounding
against hand-derived known answers, features against direct geometry."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_param_count():
    assert M.NPARAM == 9987


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, M.NPARAM)
    X = rng.normal(0, 1, (16, M.NIN))
    T = rng.normal(0, 1, (16, 3))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    L, g = M.loss_and_grad(w, X, T)
    assert abs(L - M.loss(w, X, T)) < 1e-12
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [0, 1407, 1408, 5567, 9786, 9984, 9986]])  # every block's edges
    eps = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = eps
        fd = (M.loss(w + e, X, T) - M.loss(w - e, X, T)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_minimises_quadratic():
    # f(w) = 0.5 |w - c|^2, gradient w - c; Adam with lr 0.05 reaches c
    c = np.array([1.0, -2.0, 0.5])
    w, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    for step in range(1, 2001):
        w, m, v = M.adam(w, w - c, m, v, step, 0.05)
    assert np.allclose(w, c, atol=1e-3)
    # first step moves every coordinate by lr against the gradient sign
    w1, _, _ = M.adam(np.zeros(3), np.array([3.0, -0.1, 7.0]), np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(w1, [-0.01, 0.01, -0.01], atol=1e-9)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.140625, 1e-40], np.float32)
    want = np.array([1.0, 1.0, 1.015625, -2.5, 3.140625, 0.0], np.float32)
    got = M.bf16(x)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5]) < 1e-38


def test_features_geometry(sdlib):
    from stardust_amd import hulls as H
    h = H.generate_hulls(5, 40)
    pairs = np.array([[0, 1], [5, 7], [10, 2]], np.int32)
    f = M.features(h, pairs)
    for k, (a, b) in enumerate(pairs):
        # relative position rotated back to the world equals the centre difference
        s = M.descriptor(h, a)[4] + M.descriptor(h, b)[4]
        back = M.quat_mat(h["quat4"][a]) @ (f[k, :3] * s)
        assert np.allclose(back, h["pos4"][b, :3] - h["pos4"][a, :3], atol=1e-6)
        # relative quaternion composes a's rotation into b's
        qr = f[k, 3:7]
        assert np.allclose(M.quat_mat(h["quat4"][a]) @ M.quat_mat(qr), M.quat_mat(h["quat4"][b]), atol=1e-6)
        assert np.isclose(f[k, 21], np.linalg.norm(f[k, :3]))
