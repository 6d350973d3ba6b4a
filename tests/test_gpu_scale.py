"""The scale configuration BDC_(32,16): 2^32 inputs (p and the entropy term
32 GiB each in HBM), half length 16 (the quadrant layout of the entropy
kernel, counts up to C(16,8) = 12870).  Setup takes about six minutes (the
entropy term over 2^30 representatives x 2^16 outputs); the kernel's progress
lines (one per sixteenth of the rows) are passed through to the terminal.

Checks, with the full arrays never copied: the entropy term at sampled inputs
against the oracle's single-input sums; one step from the uniform p against
the closed form p_1(x) = 2^-n exp(D_0(x)) / Z_0 (q_0 is uniform); the
split-FP16 steps against the FP64 steps from the same Dirichlet(0.5), seed 1
p; and D(x||q) of the FP64 certificate against the oracle given the engine's
ln q.
"""
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LN2 = math.log(2.0)


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(1200)  # setup alone takes about six minutes
def test_scale_bdc32_16(oracle, capfd):
    import bdc_gpu as bdc
    n, k = 32, 16
    C = math.comb(n, k)
    rng = np.random.default_rng(11)
    xs = np.array([0x0000FFFF, 0x55555555, 0x0F0F0F0F, 0x00000001] +
                  [int(v) for v in rng.integers(0, 1 << 32, size=6, dtype=np.uint64)], dtype=np.uint32)
    H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
    D0 = H / C - math.log(C) + k * LN2
    os.environ["BDC_VERBOSE"] = "1"
    say(f"scale n={n} k={k}: create")
    with capfd.disabled():  # the library's progress lines of the entropy precompute
        e = bdc.Engine(n, k)
    with e:
        assert e.num_orbits == 1073774592
        # H/C - ln C from up to 2^16 terms c ln c (sum ~6e9) over C(32,16) ~ 6e8; the
        # oracle sums in long double
        herr = np.abs(e.sample(bdc.SAMPLE_ENTROPY, xs) - (H / C - math.log(C))).max()
        say(f"scale: entropy term max error {herr:.3e} nats")
        assert herr < 1e-12
        say("scale: one FP64 step from uniform")
        e.set_p_uniform()
        st = e.iterate(1, bdc.PREC_FP64)
        lp = np.log(e.sample(bdc.SAMPLE_P, xs))
        perr = np.abs(lp + n * LN2 + st[0, 2] - D0).max()
        say(f"scale: first step max error {perr:.3e} nats")
        assert perr < 1e-10
        res = {}
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            say(f"scale: 3 steps from Dirichlet(0.5) seed 1, precision {prec}")
            e.set_p_dirichlet(0.5, 1)
            res[prec] = e.iterate(3, prec)
        np.testing.assert_allclose(res[bdc.PREC_SPLIT16], res[bdc.PREC_FP64], rtol=0, atol=1e-9)
        say("scale: certificate")
        cert = e.certify()
        assert abs(cert["sum_p"] - 1.0) < 1e-12
        logq = e.log_output()
        Dx = e.sample(bdc.SAMPLE_KL, xs)
        ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
        np.testing.assert_allclose(Dx, ref, rtol=0, atol=1e-11)
        assert cert["upper_cert"] >= ref.max() - 1e-11
