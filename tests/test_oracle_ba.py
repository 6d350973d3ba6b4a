"""The oracle's reference-method BA against a dense textbook BA and closed forms."""
import math

import numpy as np
import pytest

import dense_ba


@pytest.mark.parametrize("n,k", [(6, 3), (9, 4), (10, 5), (11, 7)])
def test_oracle_matches_dense_ba(oracle, n, k):
    W = dense_ba.channel(n, k)
    p = np.full(1 << n, 1.0 / (1 << n))
    ba = oracle.BA(n, k)
    orb = ba.orbits()
    for _ in range(6):
        pn, D, logZ, I = dense_ba.ba_step(W, p)
        mx = ba.iterate()
        assert abs(mx - (np.log2(pn) - np.log2(p)).max()) < 1e-12
        Q = ba.get_q()
        np.testing.assert_allclose(Q, pn[orb], rtol=1e-12)
        p = pn
    q = p @ W
    D = dense_ba.kl_rows(W, q)
    assert abs(ba.rate() - float(p @ D)) < 1e-12
    assert abs(ba.max_kl() - D.max()) < 1e-12
    np.testing.assert_allclose(ba.kl(), D[orb], atol=1e-12)


def test_oracle_log_den_is_log_output_law(oracle):
    n, k = 10, 6
    W = dense_ba.channel(n, k)
    p = np.full(1 << n, 1.0 / (1 << n))
    ba = oracle.BA(n, k)
    ld = ba.log_den()
    q = p @ W
    reps = [y for y in range(1 << (k - 1)) if oracle.rep(y, k) == y]
    np.testing.assert_allclose(ld[reps], np.log(q[reps]), rtol=1e-13)


def test_closed_forms(oracle):
    # C_{n,1} = 1 bit (Upper_Bounds.txt:6,42)
    ba = oracle.BA(8, 1)
    for _ in range(400):
        if ba.iterate() < 1e-12:
            break
    assert abs(ba.max_kl() / math.log(2) - 1.0) < 1e-6
    assert ba.rate() / math.log(2) <= 1.0 + 1e-12


def test_monotone_and_sandwich(oracle):
    n, k = 12, 6
    W = dense_ba.channel(n, k)
    p = np.full(1 << n, 1.0 / (1 << n))
    prevI = -1.0
    for _ in range(12):
        pn, D, logZ, I = dense_ba.ba_step(W, p)
        assert I >= prevI - 1e-13
        # log Z_t <= I(p_{t+1}) <= C <= max_x D(x||q_t)
        q1 = pn @ W
        I1 = float(pn @ dense_ba.kl_rows(W, q1))
        assert logZ <= I1 + 1e-12 and I1 <= D.max() + 1e-12
        prevI, p = I, pn
