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
    # C_{n,n} = n bits (Upper_Bounds.txt:34,72): the channel is the identity, so
    # from the uniform p one step leaves p uniform with mx = 0, and I(p) and
    # max_x D(x||q) are both n bits
    for n in (7, 10):
        ba = oracle.BA(n, n)
        assert abs(ba.iterate()) < 1e-12
        assert abs(ba.rate() / math.log(2) - n) < 1e-12
        assert abs(ba.max_kl() / math.log(2) - n) < 1e-12


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


@pytest.mark.parametrize("n,k", [(10, 5), (11, 4), (12, 11)])
def test_sample_pass_checksum(oracle, n, k):
    """or_ctx_sample_pass (the CPU-baseline workload: both reference passes over
    a slice of orbits) returns sum_i [ln Q_i + D(x_i||q_slice)] over the slice,
    q_slice being the output law of the slice alone.  Over all orbits that is
    the per-orbit log W of the reference step (BAA.cu:401-411, 487-530) summed,
    checked against a dense textbook BA; over a slice, against the dense sums
    restricted to the slice's inputs."""
    W = dense_ba.channel(n, k)
    ba = oracle.BA(n, k)
    for _ in range(2):
        ba.iterate()
    Q, orb = ba.get_q(), ba.orbits()
    sizes = np.array([oracle.orbit_size(int(x), n) for x in orb])
    rep = np.array([oracle.rep(x, n) for x in range(1 << n)])
    p = Q[np.searchsorted(orb, rep)]  # per-string probability of every input
    q = p @ W
    full = ba.sample_pass(0, ba.num_orbits)
    ref = float(np.sum(np.log(Q) + dense_ba.kl_rows(W, q)[orb]))
    assert abs(full - ref) < 1e-9 * max(1.0, abs(ref))
    a, b = ba.num_orbits // 3, ba.num_orbits // 3 + 50
    # q of the slice: its orbits' strings with their orbit weights
    in_slice = np.isin(rep, orb[a:b])
    qs = (p * in_slice) @ W
    ref_s = float(np.sum(np.log(Q[a:b]) + dense_ba.kl_rows(W, qs)[orb[a:b]]))
    assert abs(ba.sample_pass(a, b - a) - ref_s) < 1e-9 * max(1.0, abs(ref_s))
    assert (sizes > 0).all()
