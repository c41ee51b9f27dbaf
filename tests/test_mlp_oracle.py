"""Pins of the proposal-network oracle (oracle/mlp_oracle.py) without a GPU:
central finite differences of the float64 loss check every gradient block;
the forward pass is checked against closed forms."""
import numpy as np
from scipy.stats import norm

import mlp_oracle as mo


def _random_batch(rng, B=5, Kmax=6):
    x = rng.random((B, mo.NIN))
    tau = rng.uniform(0.2, 1.0, B)
    la = np.log(rng.uniform(0.1, 2.0, B))
    cls, lab = [], []
    for _ in range(B):
        K = int(rng.integers(1, Kmax + 1))
        c = np.zeros((K, mo.CLS_STRIDE))
        c[:, :20] = rng.normal(size=(K, 20)) - 3.0
        c[:, 20] = np.log(rng.integers(1, 9, K))
        cls.append(c)
        lab.append(int(rng.integers(0, K + 1)))
    return x, tau, la, cls, lab


def _theta(rng, scale=0.3):
    return rng.normal(0, scale, mo.NPARAM)


def test_gelu_matches_normal_cdf():
    z = np.linspace(-6, 6, 41)
    np.testing.assert_allclose(mo.gelu(z), z * norm.cdf(z), rtol=1e-13, atol=1e-300)
    h = 1e-5
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h),
                               rtol=1e-8, atol=1e-10)


def test_zero_network_gives_crp_weights():
    rng = np.random.default_rng(1)
    x, tau, la, cls, lab = _random_batch(rng)
    logits, _ = mo.forward(np.zeros(mo.NPARAM), x, tau, la, cls)
    for e, lg in enumerate(logits):
        w = np.concatenate([np.exp(cls[e][:, 20]), [np.exp(la[e])]])
        np.testing.assert_allclose(np.exp(mo.log_softmax(lg)), w / w.sum(), rtol=1e-13)


def test_features_mean_profile():
    rng = np.random.default_rng(2)
    lg = rng.normal(size=(3, 20))
    sizes = np.array([1, 2, 5])
    f = rng.dirichlet(np.ones(20))
    x, cls = mo.features(f, lg, sizes)
    pi = np.exp(lg) / np.exp(lg).sum(1, keepdims=True)
    np.testing.assert_allclose(x[20:], (sizes[:, None] * pi).sum(0) / 8.0, rtol=1e-13)
    np.testing.assert_allclose(np.exp(cls[:, :20]).sum(1), 1.0, rtol=1e-13)
    np.testing.assert_allclose(cls[:, 20], np.log(sizes), rtol=1e-13)


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    x, tau, la, cls, lab = _random_batch(rng)
    th = _theta(rng)
    _, g = mo.loss_grad(th, x, tau, la, cls, lab)
    # a sample of coordinates from every parameter block
    blocks = [(mo.W1, mo.B1), (mo.B1, mo.W2), (mo.W2, mo.B2), (mo.B2, mo.U), (mo.U, mo.UB),
              (mo.UB, mo.V), (mo.V, mo.BN), (mo.BN, mo.NPARAM)]
    h = 1e-6
    for a, b in blocks:
        for i in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            tp, tm = th.copy(), th.copy()
            tp[i] += h
            tm[i] -= h
            fd = (mo.loss_grad(tp, x, tau, la, cls, lab)[0] -
                  mo.loss_grad(tm, x, tau, la, cls, lab)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (a, i, fd, g[i])


def test_adam_quadratic():
    # minimising 0.5 * |theta|^2: the first Adam step moves each coordinate by
    # exactly lr * sign(g) (bias-corrected m/sqrt(v) = g/|g|)
    th = np.array([3.0, -2.0, 0.5])
    th2, m, v = mo.adam(th, np.zeros(3), np.zeros(3), th.copy(), 1, 0.1, eps=0.0)
    np.testing.assert_allclose(th2, th - 0.1 * np.sign(th), rtol=1e-14)
    for s in range(2, 400):
        th2, m, v = mo.adam(th2, m, v, th2.copy(), s, 0.05)
    assert np.abs(th2).max() < 0.05
