"""GPU parity: libbdc.so (through the C ABI) against the CPU oracle.

Counts and orbits are bit-exact; FP64 bounds agree with the oracle to 1e-10
bits (the FP64 tolerance of the metric); the split-FP16 path to 1e-9 bits.
"""
import hashlib
import json
import math
import os
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


@pytest.fixture(scope="module")
def golden():
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cache_sha256.json")
    with open(here) as f:
        return json.load(f)["sha256"]


def test_caches_match_reference_generator(bdc, golden):
    for key, digest in sorted(golden.items()):
        n, k = map(int, key.split("_"))
        say(f"cache n={n} k={k}")
        assert hashlib.sha256(bdc.generate_cache(n, k).tobytes()).hexdigest() == digest, key


def test_caches_match_oracle_beyond_fixtures(bdc, oracle):
    for n, k in [(15, 7), (16, 3)]:
        say(f"cache n={n} k={k}")
        assert np.array_equal(bdc.generate_cache(n, k), oracle.cache(n, k))


def test_write_caches_byte_identical(bdc, golden, tmp_path):
    say("write_caches 9")
    bdc.write_caches(9, str(tmp_path))
    for n in range(1, 10):
        for k in range(n):
            with open(tmp_path / f"cache_{n}_{k}", "rb") as f:
                assert hashlib.sha256(f.read()).hexdigest() == golden[f"{n}_{k}"]


@pytest.mark.parametrize("n,k", [(16, 8), (13, 6), (9, 4)])
def test_orbits(bdc, oracle, n, k):
    say(f"orbits n={n} k={k}")
    with bdc.Engine(n, k) as e:
        assert e.num_orbits == len(oracle.orbit_list(n))
        assert np.array_equal(e.orbits(), oracle.orbit_list(n))


@pytest.mark.parametrize("n,k", [(10, 5), (11, 4), (12, 7), (14, 7)])
def test_entropy_term(bdc, n, k):
    say(f"entropy n={n} k={k}")
    W = dense_ba.channel(n, k)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(W > 0, W * np.log(np.where(W > 0, W, 1.0)), 0.0).sum(axis=1)
    with bdc.Engine(n, k) as e:
        np.testing.assert_allclose(e.entropy_term(), ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k", [(12, 7), (11, 6), (13, 4)])
def test_entropy_term_quadrant_layout(bdc, monkeypatch, n, k):
    """The entropy kernel's quadrant layout (row sets restricted to one
    output quadrant, used when they do not fit LDS at half length 16) forced
    at small n, against the dense channel."""
    say(f"entropy quadrants n={n} k={k}")
    monkeypatch.setenv("BDC_ENTROPY_QUADRANTS", "1")
    W = dense_ba.channel(n, k)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(W > 0, W * np.log(np.where(W > 0, W, 1.0)), 0.0).sum(axis=1)
    with bdc.Engine(n, k) as e:
        np.testing.assert_allclose(e.entropy_term(), ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k", [(14, 7), (20, 10), (21, 10), (24, 12), (25, 9)])
@pytest.mark.parametrize("quadrants", ["0", "1"])
def test_entropy_grouped_row_sets_bitwise(bdc, monkeypatch, n, k, quadrants):
    """The x2 row sets built 8 entries per thread with packed u16
    multiply-adds from staged half tables hold the same integer counts as the
    entry-by-entry build, so the entropy term is bitwise identical (both
    layouts; half lengths 10-12 reach the 16-byte split loads)."""
    say(f"entropy grouped row sets n={n} k={k} quadrants={quadrants}")
    monkeypatch.setenv("BDC_ENTROPY_QUADRANTS", quadrants)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("BDC_ENTROPY_GROUPED", flag)
        with bdc.Engine(n, k) as e:
            out[flag] = e.entropy_term()
    assert np.array_equal(out["0"], out["1"])
    assert np.isfinite(out["1"]).all()


# (18, 9) reaches the count-streaming stage kernel (2^6-column count factor)
@pytest.mark.parametrize("n,k", [(8, 4), (10, 5), (11, 6), (12, 7), (12, 3), (18, 9)])
def test_output_law_general_weights(bdc, oracle, n, k):
    say(f"output law n={n} k={k}")
    rng = np.random.default_rng(n * 100 + k)
    w = rng.random(1 << n)
    with bdc.Engine(n, k) as e:
        q = e.output_law(w)
    np.testing.assert_allclose(q, oracle.full_output_law(n, k, w), rtol=1e-13)


@pytest.mark.parametrize("n,k", [(10, 5), (11, 5), (12, 8), (18, 10)])
def test_kl_general_q(bdc, oracle, n, k):
    say(f"kl n={n} k={k}")
    rng = np.random.default_rng(7 + n)
    q = rng.random(1 << k) + 0.05
    q /= q.sum()
    with bdc.Engine(n, k) as e:
        d = e.kl(np.log(q))
    np.testing.assert_allclose(d, oracle.full_kl(n, k, np.log(q)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k", [(14, 1), (15, 1), (14, 2), (15, 2)])
def test_small_k_matches_oracle(bdc, oracle, n, k):
    """The narrowest output spaces (2 and 4 outputs), where the stage and
    entropy kernels run their smallest levels: per-step mx, p and both bounds
    against the oracle at both precisions, and for k = 1 the closed form
    C_{n,1} = 1 bit (Upper_Bounds.txt:6,42) between I(p) and max_x D(x||q)."""
    say(f"small k n={n} k={k}")
    for prec, tol in ((bdc.PREC_FP64, 1e-10), (bdc.PREC_SPLIT16, 1e-9)):
        ref = oracle.BA(n, k)
        with bdc.Engine(n, k) as e:
            e.set_p_uniform()
            stats = e.iterate(12, prec)
            for t in range(12):
                mx = ref.iterate()
                assert abs(stats[t, 0] - mx) < tol, (prec, t, stats[t, 0], mx)
            np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=tol)
            cert = e.certify()
        assert abs(cert["lower"] - ref.rate()) / LN2 < tol
        assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < tol
        if k == 1:
            assert cert["lower"] / LN2 <= 1.0 + 1e-12 and cert["upper_cert"] / LN2 >= 1.0 - 1e-12


@pytest.mark.parametrize("n,k,iters", [(16, 8, 25), (13, 6, 15), (12, 6, 15), (18, 9, 4)])
def test_ba_fp64_matches_oracle(bdc, oracle, n, k, iters):
    say(f"ba fp64 n={n} k={k} iters={iters}")
    ref = oracle.BA(n, k)
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        stats = e.iterate(iters, bdc.PREC_FP64)
        assert stats.shape == (iters, 4)
        for t in range(iters):
            mx = ref.iterate()
            assert abs(stats[t, 0] - mx) < 1e-10, (t, stats[t, 0], mx)
        np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=1e-10)
        cert = e.certify()
        assert abs(cert["lower"] - ref.rate()) / LN2 < 1e-10
        assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < 1e-10
        assert abs(cert["sum_p"] - 1.0) < 1e-12
        # sandwich: I(p_t) <= max_x D_t, and I nondecreasing
        assert (stats[:, 3] <= stats[:, 1] + 1e-12).all()
        assert (np.diff(stats[:, 3]) >= -1e-12).all()


@pytest.mark.parametrize("n,k,precs", [(13, 6, (0,)), (14, 7, (0, 1)), (15, 7, (0, 1))])
def test_step_statistics_match_oracle(bdc, oracle, n, k, precs):
    """Every column of the per-step record -- mx, max_x D, ln sum_x p e^D and
    I(p) -- against the oracle's D(x||q_t) over the orbits before each of its
    steps.  max_x D and ln Z move with a constant shift of D (the update does
    not), so this pins the centring of ln q where W is formed from ln q itself
    (odd n) and where it is a permutation of V (even n)."""
    say(f"step statistics n={n} k={k}")
    iters = 6
    ref = oracle.BA(n, k)
    reps = ref.orbits()
    size = np.array([oracle.orbit_size(int(x), n) for x in reps], dtype=np.float64)
    want = []
    for _ in range(iters):
        w, d = size * ref.get_q(), ref.kl()
        want.append((0.0, d.max(), math.log(np.sum(w * np.exp(d))), np.sum(w * d)))
        mx = ref.iterate()
        want[-1] = (mx,) + want[-1][1:]
    want = np.array(want)
    with bdc.Engine(n, k) as e:
        for prec in precs:
            e.set_p_uniform()
            st = e.iterate(iters, prec)
            tol = 1e-10 if prec == bdc.PREC_FP64 else 1e-9
            np.testing.assert_allclose(st, want, rtol=0, atol=tol, err_msg=f"precision {prec}")


@pytest.mark.parametrize("k", [4, 6, 8, 10, 12, 14])
def test_ba_dirichlet_k_sweep(bdc, oracle, k):
    """The k-sweep workload (Dirichlet(1) init with seed 0 over orbits) at
    n = 16: both precisions against the oracle started from the same bytes."""
    n, iters = 16, 12
    say(f"dirichlet sweep n={n} k={k}")
    with bdc.Engine(n, k) as e:
        e.set_p_dirichlet(1.0, 0)
        p0 = e.get_p_orbits()
        ref = oracle.BA(n, k)
        ref.set_q(p0)
        mx_ref = np.array([ref.iterate() for _ in range(iters)])
        lb_ref, ub_ref = ref.rate() / LN2, ref.max_kl() / LN2
        for prec, tol in ((bdc.PREC_FP64, 1e-10), (bdc.PREC_SPLIT16, 1e-9)):
            e.set_p_orbits(p0)
            stats = e.iterate(iters, prec)
            np.testing.assert_allclose(stats[:, 0], mx_ref, rtol=0, atol=tol)
            cert = e.certify()
            assert abs(cert["lower"] / LN2 - lb_ref) < tol
            assert abs(cert["upper_cert"] / LN2 - ub_ref) < tol
            assert abs(cert["upper_ref"] / LN2 - (lb_ref + mx_ref[-1])) < tol


def test_ba_dirichlet_n24_matches_oracle(bdc, oracle):
    """One FP64 step of BDC_(24,8) from Dirichlet(1), seed 0, against the
    oracle's reference-method passes (about 4 M orbits on the host cores)."""
    n, k = 24, 8
    say(f"dirichlet n={n} k={k}")
    with bdc.Engine(n, k) as e:
        e.set_p_dirichlet(1.0, 0)
        p0 = e.get_p_orbits()
        stats = e.iterate(1, bdc.PREC_FP64)
        cert = e.certify()
        p1 = e.get_p_orbits()
    ref = oracle.BA(n, k)
    ref.set_q(p0)
    mx = ref.iterate()
    assert abs(stats[0, 0] - mx) < 1e-10
    np.testing.assert_allclose(p1, ref.get_q(), rtol=1e-10)
    assert abs(cert["lower"] - ref.rate()) / LN2 < 1e-10
    assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < 1e-10
    ref.close()


@pytest.mark.parametrize("n,k", [(16, 8), (15, 9), (18, 8), (20, 6), (14, 13)])
def test_factored_stage_matches_count_matrices(bdc, oracle, n, k):
    """Stage contractions as powers of the single-deletion operator (the
    default) against the count-matrix kernels (BDC_STAGE_FACTORED=0), through
    the output law and the cross term, and both against the oracle."""
    say(f"factored stage n={n} k={k}")
    rng = np.random.default_rng(5)
    w = rng.random(1 << n)
    w /= w.sum()
    q = rng.random(1 << k) + 0.01
    q /= q.sum()
    res = {}
    for flag in ("1", "0"):
        os.environ["BDC_STAGE_FACTORED"] = flag
        try:
            with bdc.Engine(n, k) as e:
                res[flag] = (e.output_law(w), e.kl(np.log(q)))
        finally:
            os.environ.pop("BDC_STAGE_FACTORED", None)
    np.testing.assert_allclose(res["1"][0], res["0"][0], rtol=1e-13, atol=0)
    np.testing.assert_allclose(res["1"][1], res["0"][1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(res["1"][0], oracle.full_output_law(n, k, w), rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(res["1"][1], oracle.full_kl(n, k, np.log(q)), rtol=0, atol=1e-12)


def test_ba_tolerance_stop(bdc, oracle):
    n, k = 12, 6
    say("ba tol stop")
    ref = oracle.BA(n, k)
    ref_iters = 0
    while True:
        ref_iters += 1
        if ref.iterate() < 1e-3:
            break
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        stats = e.iterate(5000, bdc.PREC_FP64, tol=1e-3)
    assert len(stats) == ref_iters


def test_dist_roundtrip_and_dirichlet(bdc, oracle):
    n, k = 14, 7
    say("dist roundtrip")
    with bdc.Engine(n, k) as e:
        e.set_p_dirichlet(1.0, 0)
        q1 = e.get_p_orbits()
        e.set_p_dirichlet(1.0, 0)
        assert np.array_equal(q1, e.get_p_orbits())  # deterministic in the seed
        sizes = np.array([oracle.orbit_size(int(r), n) for r in e.orbits()])
        assert abs((sizes * q1).sum() - 1.0) < 1e-12
        assert (q1 > 0).all()
        e.set_p_orbits(q1)
        assert np.array_equal(e.get_p_orbits(), q1)
        with pytest.raises(bdc.BdcError):
            e.set_p_orbits(q1[:-1])


@pytest.mark.parametrize("n,k", [(10, 5), (11, 6)])
def test_dual_gradient(bdc, oracle, n, k):
    say(f"dual n={n} k={k}")
    rng = np.random.default_rng(2)
    theta = rng.normal(scale=0.5, size=1 << k)
    with bdc.Engine(n, k) as e:
        for beta in (64.0, 256.0):
            U, g = e.dual(theta, beta)
            Uo, go = oracle.full_dual(n, k, theta, beta)
            assert abs(U - Uo) < 1e-12
            np.testing.assert_allclose(g, go, rtol=0, atol=1e-12)
            h = 1e-6
            for y in rng.choice(1 << k, 16, replace=False):
                tp, tm = theta.copy(), theta.copy()
                tp[y] += h
                tm[y] -= h
                fd = (e.dual(tp, beta, want_grad=False) - e.dual(tm, beta, want_grad=False)) / (2 * h)
                assert abs(fd - g[y]) < 1e-7


# ---------------------------------------------------------------- split-FP16
@pytest.mark.parametrize("n,k,iters", [(16, 8, 25), (14, 7, 20), (18, 9, 4)])
def test_ba_split_matches_oracle(bdc, oracle, n, k, iters):
    say(f"ba split n={n} k={k} iters={iters}")
    ref = oracle.BA(n, k)
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        stats = e.iterate(iters, bdc.PREC_SPLIT16)
        for t in range(iters):
            mx = ref.iterate()
            assert abs(stats[t, 0] - mx) < 1e-9, (t, stats[t, 0], mx)
        np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=1e-9)
        cert = e.certify()
        assert abs(cert["lower"] - ref.rate()) / LN2 < 1e-9
        assert abs(cert["upper_cert"] - ref.max_kl()) / LN2 < 1e-9


@pytest.mark.parametrize("n,k", [(14, 7), (15, 6), (16, 9)])
def test_split_building_blocks(bdc, oracle, n, k):
    say(f"split blocks n={n} k={k}")
    rng = np.random.default_rng(n + 3 * k)
    w = rng.random(1 << n)
    q = rng.random(1 << k) + 0.05
    q /= q.sum()
    with bdc.Engine(n, k) as e:
        q16 = e.output_law(w, bdc.PREC_SPLIT16)
        q64 = e.output_law(w, bdc.PREC_FP64)
        np.testing.assert_allclose(q16, q64, rtol=1e-10)
        d16 = e.kl(np.log(q), bdc.PREC_SPLIT16)
        d64 = e.kl(np.log(q), bdc.PREC_FP64)
        np.testing.assert_allclose(d16, d64, rtol=0, atol=1e-11)
    np.testing.assert_allclose(q64, oracle.full_output_law(n, k, w), rtol=1e-13)


@pytest.mark.parametrize("n,k,iters", [(24, 12, 6), (22, 10, 6)])
def test_split_matches_fp64_large(bdc, n, k, iters):
    """Larger sizes (XCD patch mapping, fused limb epilogue) against the FP64 path."""
    say(f"split vs fp64 n={n} k={k}")
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        s64 = e.iterate(iters, bdc.PREC_FP64)
        c64 = e.certify()
        e.set_p_uniform()
        s16 = e.iterate(iters, bdc.PREC_SPLIT16)
        c16 = e.certify()
    assert np.abs(s16[:, 0] - s64[:, 0]).max() < 1e-9
    assert np.abs(s16[:, 3] - s64[:, 3]).max() / LN2 < 1e-9
    assert abs(c16["upper_cert"] - c64["upper_cert"]) / LN2 < 1e-9
    assert abs(c16["lower"] - c64["lower"]) / LN2 < 1e-9


@pytest.mark.parametrize("n,k,iters", [(16, 8, 8), (20, 10, 5)])
def test_half_space_matches_full_space(bdc, monkeypatch, n, k, iters):
    """The BA step on the half space x1 < N1/2 (complement images implied)
    against the same step over all 2^n inputs, both precisions."""
    say(f"half vs full space n={n} k={k}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_HALF_SPACE", mode)
        with bdc.Engine(n, k) as e:
            for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
                e.set_p_uniform()
                st = e.iterate(iters, prec)
                out[(mode, prec)] = (st, e.certify(), e.get_p_orbits())
    for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
        (sh, ch, qh), (sf, cf, qf) = out[("1", prec)], out[("0", prec)]
        np.testing.assert_allclose(sh, sf, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(qh, qf, rtol=1e-11)
        assert abs(ch["upper_cert"] - cf["upper_cert"]) < 1e-12
        assert abs(ch["lower"] - cf["lower"]) < 1e-12
        assert abs(ch["sum_p"] - 1.0) < 1e-12


@pytest.mark.parametrize("n,k,beta", [(10, 5, 64.0), (12, 6, 256.0)])
def test_dual_gradcheck_torch(bdc, n, k, beta):
    """torch.autograd.gradcheck of the adjoint through the autograd wrapper."""
    import torch
    import bdc_torch
    say(f"dual gradcheck n={n} k={k} beta={beta}")
    theta = torch.tensor(np.random.default_rng(2).normal(scale=0.5, size=1 << k), dtype=torch.float64,
                         requires_grad=True)
    with bdc.Engine(n, k) as e:
        assert torch.autograd.gradcheck(lambda t: bdc_torch.dual_bound(t, e, beta), (theta,), eps=1e-6, atol=1e-7,
                                        rtol=1e-5)


def test_dual_value_with_gradient(bdc):
    """U_beta is the same with and without the gradient pass at a size where
    the stage contractions split K (the gradient's output law reuses the
    split-K scratch)."""
    say("dual value with/without gradient n=18 k=10")
    theta = np.random.default_rng(5).normal(scale=0.3, size=1 << 10)
    with bdc.Engine(18, 10) as e:
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            u_g, _ = e.dual(theta, 256.0, prec)
            u = e.dual(theta, 256.0, prec, want_grad=False)
            assert abs(u_g - u) < 1e-12 * max(1.0, abs(u)), (prec, u_g, u)


def _output_orbits(k):
    """the four images (y, ~y, rev y, rev ~y) of every output y"""
    ny = 1 << k
    y = np.arange(ny)
    rev = np.array([int(format(v, f"0{k}b")[::-1], 2) for v in range(ny)])
    comp = y ^ (ny - 1)
    return np.stack([y, comp, rev, rev[comp]])


def _symmetric_theta(k, seed):
    """theta with one value per complement/reversal orbit of the outputs"""
    imgs = _output_orbits(k)
    vals = np.random.default_rng(seed).normal(scale=0.5, size=1 << k)
    return vals[imgs.min(axis=0)]


def _is_symmetric(theta, k):
    imgs = _output_orbits(k)
    return all(np.array_equal(theta, theta[m]) for m in imgs[1:])


@pytest.mark.parametrize("n,k", [(14, 7), (16, 8)])
def test_dual_symmetric_half_space(bdc, oracle, monkeypatch, n, k):
    """A complement/reversal-symmetric theta takes the half-space path (the BA
    step's mirror rule over x1 < N1/2); U and dU/dtheta against the oracle's
    full-space sums, and the refinement against the full-space refinement."""
    say(f"dual symmetric n={n} k={k}")
    theta = _symmetric_theta(k, 4)
    assert _is_symmetric(theta, k)
    with bdc.Engine(n, k) as e:
        for prec, tol in ((bdc.PREC_FP64, 1e-12), (bdc.PREC_SPLIT16, 1e-9)):
            U, g = e.dual(theta, 256.0, prec)
            Uo, go = oracle.full_dual(n, k, theta, 256.0)
            assert abs(U - Uo) < tol
            np.testing.assert_allclose(g, go, rtol=0, atol=tol)
        th_half, u_half = e.refine_dual(theta, 256.0, 6, 1e-2)
        assert _is_symmetric(th_half, k)  # the orbit-averaged gradient keeps theta symmetric
        monkeypatch.setenv("BDC_HALF_SPACE", "0")
        th_full, u_full = e.refine_dual(theta, 256.0, 6, 1e-2)
        np.testing.assert_allclose(th_half, th_full, rtol=0, atol=1e-10)
        np.testing.assert_allclose(u_half, u_full, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,k", [(16, 10)])
def test_dual_refinement_beats_ba_per_pass(bdc, oracle, n, k):
    """From theta = ln q after 50 split BA steps from the uniform p, 150 Adam
    steps on U_beta (beta = 4096, lr 1e-3) certify a lower max_x D(x || q) than
    150 more BA steps from the same p (each step of either is one pass over
    the inputs; dense CPU runs of the same workload give a margin of 5e-3
    bits).  The refined bound is checked against the oracle's D(x || q)."""
    say(f"dual vs BA per pass n={n} k={k}")
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        e.iterate(50, bdc.PREC_SPLIT16)
        theta0 = e.log_output()
        th, uh = e.refine_dual(theta0, 4096.0, 150, 1e-3, bdc.PREC_SPLIT16)
        ub_dual = e.certify_q(th)
        e.iterate(150, bdc.PREC_SPLIT16)  # the refinement leaves p as it was
        ub_ba = e.certify()["upper_cert"]
    assert uh[-1] < uh[0]
    q = np.exp(th - th.max())
    D = oracle.full_kl(n, k, np.log(q / q.sum()))
    assert abs(D.max() - ub_dual) < 1e-11
    assert ub_dual < ub_ba - 1e-3 * LN2, (ub_dual / LN2, ub_ba / LN2)


@pytest.mark.parametrize("n,k", [(16, 8), (20, 10)])
def test_stopping_rule_batches_match_single_steps(bdc, n, k):
    """With a stopping rule (main_BAA.cu:73: the loop runs while mx >= a)
    bdc_iterate runs steps in batches from a snapshot of the step state and
    reruns a batch that crosses the stopping point.  It must stop at the same
    step, with bitwise the same statistics and p, as one call per step and as
    the same number of steps without the rule."""
    say(f"stopping rule n={n} k={k}")
    with bdc.Engine(n, k) as e:
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            e.set_p_dirichlet(1.0, 5)
            mx = e.iterate(90, prec)[:, 0]
            # the first new running minimum after step 40 that is not the last
            # step of a batch of 32 (so the batch is rerun up to it): tol
            # between it and the minimum before it
            run_min = np.minimum.accumulate(mx)
            stop = next((i for i in range(40, 90) if i % 32 != 31 and mx[i] < run_min[i - 1]), None)
            assert stop is not None, mx
            tol = 0.5 * (mx[stop] + run_min[stop - 1])
            e.set_p_dirichlet(1.0, 5)
            batched = e.iterate(90, prec, tol=tol)
            p_batched = e.get_p_orbits()
            e.set_p_dirichlet(1.0, 5)
            single = np.concatenate([e.iterate(1, prec, tol=tol) for _ in range(stop + 1)])
            p_single = e.get_p_orbits()
            e.set_p_dirichlet(1.0, 5)
            free = e.iterate(stop + 1, prec)
            p_free = e.get_p_orbits()
            assert len(batched) == stop + 1, prec
            assert np.array_equal(batched, single) and np.array_equal(batched, free), prec
            assert np.array_equal(p_batched, p_single) and np.array_equal(p_batched, p_free), prec


@pytest.mark.parametrize("n,k", [(16, 8), (20, 10)])
def test_runs_are_bitwise_deterministic(bdc, n, k):
    """No reduction uses atomics (block partials are combined in a fixed order),
    so two runs of the same steps give bitwise identical statistics, p and
    certificate at both precisions; the reference's FP64 atomicAdd into the
    output law (BAA.cu:477-480) does not have this property."""
    say(f"determinism n={n} k={k}")
    with bdc.Engine(n, k) as e:
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            runs = []
            for _ in range(2):
                e.set_p_dirichlet(1.0, 3)
                st = e.iterate(8, prec)
                runs.append((st, e.get_p_orbits(), e.certify()))
            (s0, p0, c0), (s1, p1, c1) = runs
            assert np.array_equal(s0, s1) and np.array_equal(p0, p1), prec
            assert c0 == c1, prec


@pytest.mark.parametrize("n,k", [(24, 12), (20, 10), (22, 14)])
def test_fp64_aligned_kernels_match_generic(bdc, monkeypatch, n, k):
    """The aligned FP64 product kernels (16-byte loads, two LDS stages) against
    the generic strided FP64 kernel (BDC_GEMM64=legacy): the same BA steps,
    certificate and output law to rounding (the sums run in another order)."""
    say(f"fp64 aligned vs generic n={n} k={k}")
    rng = np.random.default_rng(n + k)
    w = rng.random(1 << n)
    out = {}
    for mode in ("aligned", "legacy"):
        monkeypatch.setenv("BDC_GEMM64", mode)
        with bdc.Engine(n, k) as e:
            e.set_p_dirichlet(1.0, 7)
            st = e.iterate(6, bdc.PREC_FP64)
            out[mode] = (st, e.certify(), e.output_law(w, bdc.PREC_FP64))
    (sa, ca, qa), (sl, cl, ql) = out["aligned"], out["legacy"]
    np.testing.assert_allclose(sa, sl, rtol=1e-13, atol=1e-14)
    assert abs(ca["upper_cert"] - cl["upper_cert"]) < 1e-13
    assert abs(ca["lower"] - cl["lower"]) < 1e-13
    np.testing.assert_allclose(qa, ql, rtol=1e-13)


@pytest.mark.parametrize("n,k", [(16, 8), (20, 10), (24, 12)])
def test_quarter_triangle_reads_match_mirrored(bdc, monkeypatch, n, k):
    """Quarter-layout split steps that update only the canonical elements, the
    output law reading each column from the triangle (the default), against
    steps that write both members of every reversal pair
    (BDC_QUARTER_TRI=0).  The output law reads the same values with the same
    chunk exponents either way, so the steps, certificate and p agree
    bitwise."""
    say(f"quarter triangle vs mirrored n={n} k={k}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_QUARTER_TRI", mode)
        with bdc.Engine(n, k) as e:
            e.set_p_dirichlet(1.0, 13)
            st = e.iterate(9, bdc.PREC_SPLIT16)
            st2 = e.iterate(3, bdc.PREC_SPLIT16)  # a second call starts from the filled half space
            out[mode] = (np.vstack([st, st2]), e.certify(), e.get_p_orbits())
    (sa, ca, pa), (sb, cb, pb) = out["1"], out["0"]
    assert np.array_equal(sa, sb)
    assert ca == cb
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("n,k,bits", [(16, 8, 9), (20, 10, 8), (24, 12, 7)])
def test_g1_bit_field_limbs_match_rint_chain(bdc, monkeypatch, n, k, bits):
    """The output law's five limbs of p taken as bit fields of the value
    rounded to its last limb (the default for limbs of 7 to 9 bits, forced here
    through the limb width) against the chain of rint steps
    (BDC_G1_FIELDS=0).  Both represent the same rounded values (ties to even),
    so the output law and the steps agree to the FP32 rounding of the
    accumulator the last two limbs share; signed weights exercise negative
    leading limbs, and both agree with FP64 to 2^-40 of the law of |w|."""
    say(f"G1 bit-field limbs n={n} k={k} bits={bits}")
    monkeypatch.setenv("BDC_SPLIT_G1_BITS", str(bits))
    rng = np.random.default_rng(n + bits)
    w = rng.standard_normal(1 << n)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_G1_FIELDS", mode)
        with bdc.Engine(n, k) as e:
            cfg = e.split_config()
            assert (cfg["g1_bits"], cfg["g1_limbs"]) == (bits, 5)
            law = (e.output_law(w, bdc.PREC_SPLIT16), e.output_law(w), e.output_law(np.abs(w)))
            e.set_p_dirichlet(1.0, 5)
            st = e.iterate(6, bdc.PREC_SPLIT16)
            out[mode] = (law, st, e.certify())
    (la, sa, ca), (lb, sb, cb) = out["1"], out["0"]
    for q16, q64, qabs in (la, lb):
        assert np.all(np.abs(q16 - q64) <= 2.0**-40 * qabs)
    assert np.all(np.abs(la[0] - lb[0]) <= 2.0**-44 * la[2])
    np.testing.assert_allclose(sa, sb, rtol=1e-12, atol=1e-14)
    assert abs(ca["upper_cert"] - cb["upper_cert"]) < 1e-12
    assert abs(ca["lower"] - cb["lower"]) < 1e-12


@pytest.mark.parametrize("n,k", [(20, 10), (24, 12)])
def test_g1_uniform_loops_bitwise(bdc, monkeypatch, n, k):
    """The output law's copies of its loop for the stages wholly inside the
    triangle (read in place) and wholly outside it (read from the partners as
    16-byte pairs), without the per-element source select, against the general
    loop alone (BDC_G1_UNIFORM_LOOPS=0): the same values in the same order, so
    the steps agree bitwise.  (20,10) has the transposed copy only; (24,12) all
    three, and a second K slab."""
    say(f"G1 uniform loops n={n} k={k}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_G1_UNIFORM_LOOPS", mode)
        with bdc.Engine(n, k) as e:
            e.set_p_dirichlet(1.0, 17)
            st = e.iterate(5, bdc.PREC_SPLIT16)
            out[mode] = (st, e.certify(), e.get_p_orbits())
    (sa, ca, pa), (sb, cb, pb) = out["1"], out["0"]
    assert np.array_equal(sa, sb)
    assert ca == cb
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("n,k", [(16, 8), (24, 12)])
def test_vw_limbs_one_launch_bitwise(bdc, monkeypatch, n, k):
    """W formed from V under the mirror rule together with the V and W limbs in
    one launch (the default) against W from V, then the limbs of V and of W in
    their own launches (BDC_VW_FUSED=0): the same arithmetic on the same
    values, so split steps agree bitwise."""
    say(f"V/W limbs in one launch n={n} k={k}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_VW_FUSED", mode)
        with bdc.Engine(n, k) as e:
            e.set_p_dirichlet(1.0, 23)
            st = e.iterate(5, bdc.PREC_SPLIT16)
            out[mode] = (st, e.certify(), e.get_p_orbits())
    (sa, ca, pa), (sb, cb, pb) = out["1"], out["0"]
    assert np.array_equal(sa, sb)
    assert ca == cb
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("n,k", [(16, 8), (20, 10), (24, 12)])
def test_column_order_matches_natural(bdc, monkeypatch, n, k):
    """P, H and D stored with bit-reversed columns (the reversal becomes a
    transpose) against natural order (BDC_QUARTER=0): the same steps,
    certificate, dist-file values and user-facing x-indexed arrays."""
    say(f"column order n={n} k={k}")
    rng = np.random.default_rng(n * k)
    w = rng.random(1 << n)
    q = rng.random(1 << k) + 0.1
    q /= q.sum()
    xs = rng.integers(0, 1 << n, size=64, dtype=np.uint64).astype(np.uint32)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BDC_QUARTER", mode)
        with bdc.Engine(n, k) as e:
            res = []
            for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
                e.set_p_dirichlet(1.0, 11)
                st = e.iterate(7, prec)
                res.append((st, e.certify(), e.get_p_orbits(), e.sample(bdc.SAMPLE_P, xs), e.sample(bdc.SAMPLE_KL, xs)))
            res.append((e.entropy_term(), e.output_law(w), e.kl(np.log(q)), e.sample(bdc.SAMPLE_ENTROPY, xs)))
            out[mode] = res
    # every element is computed by the same arithmetic in both layouts; the block
    # partials of the step's sums (and the certificate's sums over columns) group
    # the inputs differently, and seven steps from a Dirichlet start carry those
    # rounding differences to a few 1e-12 (the oracle tests bound both to 1e-10)
    for a, b in zip(out["1"][:2], out["0"][:2]):
        np.testing.assert_allclose(a[0], b[0], rtol=1e-10, atol=1e-12)
        assert abs(a[1]["upper_cert"] - b[1]["upper_cert"]) < 1e-10
        assert abs(a[1]["lower"] - b[1]["lower"]) < 1e-10
        for i in (2, 3, 4):
            np.testing.assert_allclose(a[i], b[i], rtol=1e-10)
    for a, b in zip(out["1"][2], out["0"][2]):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-13)
