"""GPU parity of the deletion-chain path (output lengths k > 20 and k = n; the
reference enumerates the distinct subsequences of each input for any k <= n,
backend/BAA.cu:429-485, 642-724) through the C ABI.

At small sizes the chain path is forced (BDC_CHAIN_PATH=1) and checked
against the oracle's full-space sums and its reference-method BA; the
enumeration entropy kernel is forced (BDC_ENTROPY_DFS=1) and checked against
the dense channel and the dense sweep kernel.  Native k > 20 contexts run
against the oracle's BA at n = 22 and 24, k = n against the closed form
C_{n,n} = n bits (Upper_Bounds.txt:34,72), and BDC_(29,24) and BDC_(31,29) at
sampled inputs (entropy term, the closed form of the first step from uniform p,
and the certificate's D(x||q)).
"""
import math
import sys

import numpy as np
import pytest

import dense_ba

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bdc():
    import bdc_gpu
    bdc_gpu.lib()
    return bdc_gpu


def dense_entropy(n, k):
    W = dense_ba.channel(n, k)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(W > 0, W * np.log(np.where(W > 0, W, 1.0)), 0.0).sum(axis=1)


@pytest.mark.parametrize("n,k", [(12, 9), (13, 11), (12, 12), (11, 1), (10, 4), (14, 7), (13, 6)])
def test_forced_chain_building_blocks(bdc, oracle, monkeypatch, n, k):
    """output law, D(x||q) and the entropy term of the chain path against the
    oracle and the dense channel (groups: one 13-bit group, or 13 + the rest)"""
    say(f"chain blocks n={n} k={k}")
    monkeypatch.setenv("BDC_CHAIN_PATH", "1")
    rng = np.random.default_rng(n * 31 + k)
    w = rng.random(1 << n)
    q = rng.random(1 << k) + 0.05
    q /= q.sum()
    with bdc.Engine(n, k) as e:
        if n - k <= 16:
            np.testing.assert_allclose(e.entropy_term(), dense_entropy(n, k), rtol=0, atol=1e-12)
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):  # the chain path is FP64 at either setting
            np.testing.assert_allclose(e.output_law(w, prec), oracle.full_output_law(n, k, w), rtol=1e-13)
            np.testing.assert_allclose(e.kl(np.log(q), prec), oracle.full_kl(n, k, np.log(q)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k,iters", [(14, 11, 12), (16, 14, 8), (15, 15, 3), (13, 5, 10), (16, 9, 5)])
def test_forced_chain_ba_matches_oracle(bdc, oracle, monkeypatch, n, k, iters):
    """BA steps, p and both bounds of the chain path against the reference
    method, from uniform p (16 = 13 + 3 bits: two groups)"""
    say(f"chain BA n={n} k={k}")
    monkeypatch.setenv("BDC_CHAIN_PATH", "1")
    ref = oracle.BA(n, k)
    mx_ref = [ref.iterate() for _ in range(iters)]
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        st = e.iterate(iters, bdc.PREC_FP64)
        np.testing.assert_allclose(st[:, 0], mx_ref, rtol=0, atol=1e-10)
        np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=1e-10)
        cert = e.certify()
        assert abs(cert["lower"] - ref.rate()) / LN2 < 1e-10
        assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < 1e-10
        # the no-allocation entry point runs the same steps
        e.set_p_uniform()
        assert np.array_equal(e.iterate_noalloc(iters, bdc.PREC_FP64), st)


@pytest.mark.parametrize("n,k", [(16, 10), (18, 12), (17, 15), (14, 7), (16, 13), (16, 12), (18, 10), (19, 10),
                                 (20, 10), (21, 10), (20, 8), (21, 8), (22, 8), (23, 8), (24, 8)])
def test_entropy_enumeration_matches_dense_sweep(bdc, monkeypatch, n, k):
    """the distinct-subsequence entropy kernel against the dense 2^k sweep on
    the same dense-path context shape, for every number of deletions d = n - k
    from 2 to 16 (the kernel is instantiated per d)"""
    say(f"entropy enumeration n={n} k={k}")
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("BDC_ENTROPY_DFS", flag)
        with bdc.Engine(n, k) as e:
            out[flag] = e.entropy_term()
    np.testing.assert_allclose(out["1"], out["0"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,k,iters", [(22, 21, 3), (24, 22, 2)])
def test_native_large_k_matches_oracle(bdc, oracle, n, k, iters):
    """k > 20 contexts (the chain path without forcing) against the reference
    method over all orbits"""
    say(f"large k n={n} k={k}")
    ref = oracle.BA(n, k)
    mx_ref = [ref.iterate() for _ in range(iters)]
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        st = e.iterate(iters, bdc.PREC_SPLIT16)
        np.testing.assert_allclose(st[:, 0], mx_ref, rtol=0, atol=1e-10)
        np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=1e-10)
        cert = e.certify()
        assert abs(cert["lower"] - ref.rate()) / LN2 < 1e-10
        assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < 1e-10
    ref.close()


@pytest.mark.parametrize("n", [12, 20])
def test_k_equals_n_closed_form(bdc, n):
    """C_{n,n} = n bits: the channel is the identity, the uniform p is
    optimal, one step leaves it unchanged (mx = 0) and I(p) = max_x D = n bits"""
    say(f"k = n = {n}")
    with bdc.Engine(n, n) as e:
        e.set_p_uniform()
        st = e.iterate(2, bdc.PREC_FP64)
        assert np.abs(st[:, 0]).max() < 1e-12
        cert = e.certify()
        assert abs(cert["lower"] / LN2 - n) < 1e-12
        assert abs(cert["upper_cert"] / LN2 - n) < 1e-12
        assert np.abs(e.entropy_term()).max() == 0.0


def test_bdc29_24_sampled(bdc, oracle):
    """BDC_(29,24) (2^29 inputs, odd n): entropy term at sampled inputs, the
    first step from uniform against p_1(x) = 2^-n exp(D_0(x)) / Z_0 with
    D_0(x) = H(x)/C - ln C + k ln 2, and after three steps the certificate's
    D(x||q) against the oracle given the engine's ln q"""
    n, k = 29, 24
    C = math.comb(n, k)
    rng = np.random.default_rng(2924)
    xs = np.array([0, (1 << n) - 1, int("01" * 14 + "0", 2), 0x0F0F0F0F & ((1 << n) - 1)] +
                  [int(v) for v in rng.integers(0, 1 << n, size=6, dtype=np.uint64)], dtype=np.uint32)
    H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
    D0 = H / C - math.log(C) + k * LN2
    say(f"BDC_({n},{k}): create")
    with bdc.Engine(n, k) as e:
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_ENTROPY, xs), H / C - math.log(C), rtol=0, atol=1e-12)
        e.set_p_uniform()
        st = e.iterate(1, bdc.PREC_FP64)
        lp = np.log(e.sample(bdc.SAMPLE_P, xs))
        np.testing.assert_allclose(lp + n * LN2 + st[0, 2], D0, rtol=0, atol=1e-10)
        assert st[0, 1] >= D0.max() - 1e-10
        st = np.vstack([st, e.iterate(2, bdc.PREC_FP64)])
        assert (np.diff(st[:, 3]) >= -1e-12).all()
        say(f"BDC_({n},{k}): certificate")
        cert = e.certify()
        assert abs(cert["sum_p"] - 1.0) < 1e-12
        logq = e.log_output()
        ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_KL, xs), ref, rtol=0, atol=1e-11)
        assert cert["upper_cert"] >= ref.max() - 1e-11


def test_bdc31_29_sampled(bdc, oracle):
    """BDC_(31,29) (2^31 inputs; bit groups 9 + 9 + 13, where n = 29 has 8 + 8 +
    13): entropy term at sampled inputs, the first step from uniform against its
    closed form, and after two more steps the certificate's D(x||q) against the
    oracle given the engine's ln q"""
    n, k = 31, 29
    C = math.comb(n, k)
    rng = np.random.default_rng(3129)
    xs = np.array([0, (1 << n) - 1, int("01" * 15 + "0", 2), 0x0F0F0F0F & ((1 << n) - 1)] +
                  [int(v) for v in rng.integers(0, 1 << n, size=6, dtype=np.uint64)], dtype=np.uint32)
    H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
    D0 = H / C - math.log(C) + k * LN2
    say(f"BDC_({n},{k}): create")
    with bdc.Engine(n, k) as e:
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_ENTROPY, xs), H / C - math.log(C), rtol=0, atol=1e-12)
        e.set_p_uniform()
        st = e.iterate(1, bdc.PREC_FP64)
        lp = np.log(e.sample(bdc.SAMPLE_P, xs))
        np.testing.assert_allclose(lp + n * LN2 + st[0, 2], D0, rtol=0, atol=1e-10)
        assert st[0, 1] >= D0.max() - 1e-10
        st = np.vstack([st, e.iterate(2, bdc.PREC_SPLIT16)])
        assert (np.diff(st[:, 3]) >= -1e-12).all()
        say(f"BDC_({n},{k}): certificate")
        cert = e.certify()
        assert abs(cert["sum_p"] - 1.0) < 1e-12
        logq = e.log_output()
        ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_KL, xs), ref, rtol=0, atol=1e-11)
        assert cert["upper_cert"] >= ref.max() - 1e-11


@pytest.mark.parametrize("n,k", [(12, 10), (11, 11)])
def test_chain_dual_gradient(bdc, oracle, monkeypatch, n, k):
    """U_beta and its adjoint gradient on the chain path against the oracle"""
    say(f"chain dual n={n} k={k}")
    monkeypatch.setenv("BDC_CHAIN_PATH", "1")
    theta = np.random.default_rng(3).normal(scale=0.5, size=1 << k)
    with bdc.Engine(n, k) as e:
        U, g = e.dual(theta, 64.0)
        Uo, go = oracle.full_dual(n, k, theta, 64.0)
        assert abs(U - Uo) < 1e-12
        np.testing.assert_allclose(g, go, rtol=0, atol=1e-12)
