"""Dense textbook Blahut–Arimoto on the full 2^n x 2^k channel matrix (numpy).

An independent check of the oracle for small n (n <= 12): it uses no symmetry,
no split caches and no subsequence enumeration — only the definition
W(y|x) = c(x,y)/C(n,k), with c from brute force over position subsets.
"""
import itertools
import math

import numpy as np


def brute_count(x, n, y, k):
    """Occurrences of y as a subsequence of x by enumerating all C(n,k) position sets."""
    xs = [(x >> (n - 1 - i)) & 1 for i in range(n)]
    ys = [(y >> (k - 1 - i)) & 1 for i in range(k)]
    return sum(1 for pos in itertools.combinations(range(n), k) if all(xs[p] == ys[j] for j, p in enumerate(pos)))


def count_matrix(n, k):
    """c(x,y) for all x, y by a vectorised prefix DP (equal to brute_count; tested)."""
    X = np.arange(1 << n)[:, None]
    Y = np.arange(1 << k)[None, :]
    dp = [np.ones((1 << n, 1 << k), dtype=np.int64)] + [np.zeros((1 << n, 1 << k), dtype=np.int64) for _ in range(k)]
    for i in range(n):
        xb = (X >> (n - 1 - i)) & 1
        for j in range(min(i + 1, k), 0, -1):
            yb = (Y >> (k - j)) & 1
            dp[j] = dp[j] + (xb == yb) * dp[j - 1]
    return dp[k]


def channel(n, k):
    return count_matrix(n, k).astype(np.float64) / math.comb(n, k)


def kl_rows(W, q):
    """D(x||q) for every row x of W (0 log 0 = 0)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(W > 0, W * (np.log(np.where(W > 0, W, 1.0)) - np.log(q)[None, :]), 0.0)
    return t.sum(axis=1)


def ba_step(W, p):
    """One textbook BA step: returns (p_next, D, log Z, I(p)) for the current p."""
    q = p @ W
    D = kl_rows(W, q)
    w = p * np.exp(D - D.max())
    Z = w.sum()
    logZ = math.log(Z) + D.max()
    return w / Z, D, logZ, float(p @ D)
