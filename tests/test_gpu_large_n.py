"""GPU parity beyond the sizes the oracle can enumerate: odd n = 29 (half
lengths 15 / 14, the general non-mirrored path over the full input space).

The full 2^29-input arrays are not copied; values are sampled at chosen
inputs (bdc_sample_inputs) and checked against the oracle's single-input
sums.  From the reference's uniform p (backend/main_BAA.cu:64) the output law
is uniform, q_0 = 2^-k (every y is reached by C(n,k) 2^(n-k) (x, position)
pairs), so after one step p_1(x) = 2^-n exp(D_0(x)) / Z_0 with
D_0(x) = H(x)/C - ln C + k ln 2, which pins the entropy term, the update and
ln Z_0 at this size.  D(x||q_1) from the certificate is then compared with the
oracle given the engine's FP64 ln q_1, and the split-FP16 step with the FP64
one.
"""
import math
import sys

import numpy as np
import pytest

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


def sample_inputs(n, count, seed):
    rng = np.random.default_rng(seed)
    xs = [0, (1 << (n // 2)) - 1, int("01" * (n // 2) + "0" * (n % 2), 2)]
    xs += [int(v) for v in rng.integers(0, 1 << n, size=count, dtype=np.uint64)]
    return np.array(xs, dtype=np.uint32)


def test_odd_n29_fp64_and_split(bdc, oracle):
    n, k = 29, 14
    C = math.comb(n, k)
    xs = sample_inputs(n, 6, 7)
    H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
    D0 = H / C - math.log(C) + k * LN2
    say(f"large n={n} k={k}: create")
    with bdc.Engine(n, k) as e:
        # entropy term at the sampled inputs
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_ENTROPY, xs), H / C - math.log(C), rtol=0, atol=1e-12)
        results = {}
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            say(f"large n={n} k={k}: steps at precision {prec}")
            e.set_p_uniform()
            st = e.iterate(1, prec)
            # p_1(x) = 2^-n exp(D_0(x)) / Z_0
            lp = np.log(e.sample(bdc.SAMPLE_P, xs))
            tol = 1e-10 if prec == bdc.PREC_FP64 else 1e-9
            np.testing.assert_allclose(lp + n * LN2 + st[0, 2], D0, rtol=0, atol=tol)
            assert st[0, 1] >= D0.max() - tol  # max_x D_0 over all x covers the sample
            results[prec] = np.vstack([st, e.iterate(3, prec)])
        # the two precisions agree on every step's statistics (mx in bits, the rest in nats)
        np.testing.assert_allclose(results[bdc.PREC_SPLIT16], results[bdc.PREC_FP64], rtol=0, atol=5e-10)
        # three more FP64 steps, then the certificate's D(x||q) against the oracle
        say(f"large n={n} k={k}: certificate")
        e.set_p_uniform()
        st = e.iterate(3, bdc.PREC_FP64)
        assert (np.diff(st[:, 3]) >= -1e-12).all()
        cert = e.certify()
        assert abs(cert["sum_p"] - 1.0) < 1e-12
        assert cert["lower"] <= cert["upper_cert"]
        logq = e.log_output()
        Dx = e.sample(bdc.SAMPLE_KL, xs)
        ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
        np.testing.assert_allclose(Dx, ref, rtol=0, atol=1e-11)
        assert cert["upper_cert"] >= ref.max() - 1e-11
