"""Smoothed dual bound of the oracle: closed-form gradient vs central differences."""
import numpy as np

import dense_ba


def test_dual_gradient_fd(oracle):
    n, k = 8, 4
    rng = np.random.default_rng(2)
    theta = rng.normal(size=1 << k)
    for beta in (4.0, 64.0):
        U, g = oracle.full_dual(n, k, theta, beta)
        h = 1e-6
        for y in rng.choice(1 << k, 8, replace=False):
            tp, tm = theta.copy(), theta.copy()
            tp[y] += h
            tm[y] -= h
            fd = (oracle.full_dual(n, k, tp, beta, False) - oracle.full_dual(n, k, tm, beta, False)) / (2 * h)
            assert abs(fd - g[y]) < 1e-7 * max(1.0, abs(g[y]))
        # U >= max_x D and the gradient sums to zero (softmax invariance)
        q = np.exp(theta - theta.max())
        q /= q.sum()
        D = dense_ba.kl_rows(dense_ba.channel(n, k), q)
        assert U >= D.max() - 1e-12
        assert abs(g.sum()) < 1e-12
