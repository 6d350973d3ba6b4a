"""bdc_iterate_noalloc (the entry point that allocates no device memory; its
per-step statistics live in a file-scope device array) against bdc_iterate:
bitwise the same statistics and p on two calls in a row, the second with
fewer steps, at both precisions and on the full and the half input space."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,k,big,small", [(16, 8, 40, 7), (20, 10, 12, 3), (13, 6, 9, 2)])
def test_noalloc_matches_iterate(n, k, big, small):
    import bdc_gpu as bdc
    say(f"noalloc n={n} k={k}")
    with bdc.Engine(n, k) as e:
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            if prec == bdc.PREC_SPLIT16 and n < 14:
                continue  # the split path needs 128 x 128 tiles of p
            e.set_p_dirichlet(1.0, 4)
            e.iterate(1, prec)  # prepares the context (split operands, history buffer)
            for steps in (big, small):
                e.set_p_dirichlet(1.0, 4)
                ref = e.iterate(steps, prec)
                p_ref, c_ref = e.get_p_orbits(), e.certify()
                e.set_p_dirichlet(1.0, 4)
                got = e.iterate_noalloc(steps, prec)
                assert np.array_equal(got, ref), (prec, steps)
                assert np.array_equal(e.get_p_orbits(), p_ref), (prec, steps)
                assert e.certify() == c_ref, (prec, steps)


def test_noalloc_rejects():
    import bdc_gpu as bdc
    say("noalloc errors")
    with bdc.Engine(16, 8) as e:
        e.set_p_uniform()
        with pytest.raises(bdc.BdcError) as err:
            e.iterate_noalloc(5, bdc.PREC_SPLIT16)  # split operands not prepared yet
        assert err.value.code == bdc.ERR_STATE
        with pytest.raises(bdc.BdcError) as err:
            e.iterate_noalloc(4097, bdc.PREC_FP64)
        assert err.value.code == bdc.ERR_ARG
        assert e.iterate_noalloc(0, bdc.PREC_FP64).shape == (0, 4)
