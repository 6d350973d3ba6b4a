"""GPU parity at the CPU-runnable config BDC_(16,8) run to its tolerance, and at
the full sizes of the precision config BDC_(28,14) (the metric workload) and the
differentiation config BDC_(28,18), through the C ABI.

The full 2^28-input arrays are never enumerated on the CPU.  Exact references
come from the oracle's single-input sums and from weights with a known output
law:
  - from the reference's uniform p (backend/main_BAA.cu:64) the output law is
    uniform, q_0 = 2^-k, so p_1(x) = 2^-n exp(D_0(x)) / Z_0 with
    D_0(x) = H(x)/C - ln C + k ln 2 at any sampled x;
  - p = (1 - eps) uniform + eps (mass on ~1000 random orbits) has the output law
    (1 - eps) 2^-k + eps q_S, and q_S (a sum over the support's strings) is
    exact on the CPU; it pins the output law (compute_den_kernel,
    backend/BAA.cu:429-485), one BA step at sampled inputs (BAA.cu:590-639) and
    the certificate's D(x||q) (compute_KL_kernel, BAA.cu:533-571) at both
    precisions.
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


def orbit_images(x, n):
    """the strings of x's complement/reversal orbit"""
    full = (1 << n) - 1
    rev = int(format(x, f"0{n}b")[::-1], 2)
    return sorted({x, rev, x ^ full, rev ^ full})


def random_inputs(n, count, seed):
    rng = np.random.default_rng(seed)
    xs = [0, (1 << n) - 1, int("01" * (n // 2), 2), (1 << (n // 2)) - 1]
    xs += [int(v) for v in rng.integers(0, 1 << n, size=count, dtype=np.uint64)]
    return np.array(xs, dtype=np.uint32)


# ----------------------------------------------------------------- BDC_(16,8)
def test_bdc16_8_to_tolerance(bdc, oracle):
    """The CPU-runnable config: FP64 BA from the uniform p until the bound gap
    mx < 1e-6 bits (the reference's stopping rule, main_BAA.cu:73; here
    max_x D - I(p) < 1e-6 bits as well).  The first 300 steps against the
    oracle step by step; the oracle cannot follow the ~6e4 steps to the
    tolerance, so it certifies the engine's final p instead (I(p) and
    max_x D(x||q), BAA.cu:533-571, 836-913); split16 from the same start stops
    within a few steps of FP64 with bounds within 1e-9 bits."""
    n, k, tol = 16, 8, 1e-6
    say("BDC_(16,8): 300 steps against the oracle")
    ref = oracle.BA(n, k)
    mx_ref = [ref.iterate() for _ in range(300)]
    with bdc.Engine(n, k) as e:
        e.set_p_uniform()
        st = e.iterate(300, bdc.PREC_FP64)
        np.testing.assert_allclose(st[:, 0], mx_ref, rtol=0, atol=1e-10)
        np.testing.assert_allclose(e.get_p_orbits(), ref.get_q(), rtol=1e-9)
        say("BDC_(16,8): FP64 and split16 to mx < 1e-6 bits")
        runs = {}
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            e.set_p_uniform()
            steps = 0
            while steps < 400_000:
                st = e.iterate(50_000, prec, tol=tol)
                steps += len(st)
                if st[-1, 0] < tol:
                    break
            assert st[-1, 0] < tol, (prec, steps)
            cert = e.certify()
            runs[prec] = (steps, cert, e.get_p_orbits())
        (s64, c64, q64), (s16, c16, _) = runs[bdc.PREC_FP64], runs[bdc.PREC_SPLIT16]
        assert (c64["upper_cert"] - c64["lower"]) / LN2 < tol
        assert abs(s16 - s64) <= 5, (s16, s64)
        assert abs(c16["upper_cert"] - c64["upper_cert"]) / LN2 < 1e-9
        assert abs(c16["lower"] - c64["lower"]) / LN2 < 1e-9
    say("BDC_(16,8): the oracle certifies the final FP64 p")
    ref.set_q(q64)
    assert abs(ref.rate() - c64["lower"]) / LN2 < 1e-10
    assert abs(ref.max_kl() - c64["upper_cert"]) / LN2 < 1e-10
    ref.close()


# ----------------------------------------------------------------- BDC_(28,14)
@pytest.fixture(scope="module")
def e2814(bdc):
    say("BDC_(28,14): create")
    e = bdc.Engine(28, 14)
    yield e
    e.close()


def test_bdc28_14_uniform_start(bdc, oracle, e2814):
    """entropy term at sampled inputs; the first step from uniform against its
    closed form at both precisions; 50 split-FP16 steps against 50 FP64 steps
    (1e-9 bits per step on mx, I(p) and I(p) + mx); the FP64 certificate's
    D(x||q) at sampled inputs against the oracle given the engine's ln q."""
    e, n, k = e2814, 28, 14
    C = math.comb(n, k)
    xs = random_inputs(n, 8, 21)
    H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
    np.testing.assert_allclose(e.sample(bdc.SAMPLE_ENTROPY, xs), H / C - math.log(C), rtol=0, atol=1e-12)
    D0 = H / C - math.log(C) + k * LN2
    runs = {}
    for prec, tol in ((bdc.PREC_FP64, 1e-10), (bdc.PREC_SPLIT16, 1e-9)):
        say(f"BDC_(28,14): 50 steps from uniform at precision {prec}")
        e.set_p_uniform()
        st = e.iterate(1, prec)
        lp = np.log(e.sample(bdc.SAMPLE_P, xs))
        np.testing.assert_allclose(lp + n * LN2 + st[0, 2], D0, rtol=0, atol=tol)
        assert st[0, 1] >= D0.max() - tol
        runs[prec] = (np.vstack([st, e.iterate(49, prec)]), e.certify())
    (s64, c64), (s16, c16) = runs[bdc.PREC_FP64], runs[bdc.PREC_SPLIT16]
    assert np.abs(s16[:, 0] - s64[:, 0]).max() < 1e-9
    assert np.abs(s16[:, 3] - s64[:, 3]).max() / LN2 < 1e-9
    assert np.abs((s16[:, 3] / LN2 + s16[:, 0]) - (s64[:, 3] / LN2 + s64[:, 0])).max() < 1e-9
    assert abs(c16["upper_cert"] - c64["upper_cert"]) / LN2 < 1e-9
    assert abs(c16["lower"] - c64["lower"]) / LN2 < 1e-9
    # the split run's p is current: certificate at sampled inputs
    logq = e.log_output()
    Dx = e.sample(bdc.SAMPLE_KL, xs)
    ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
    np.testing.assert_allclose(Dx, ref, rtol=0, atol=1e-11)
    assert c16["upper_cert"] >= ref.max() - 1e-11
    assert abs(c16["sum_p"] - 1.0) < 1e-12


def test_bdc28_14_sparse_support(bdc, oracle, e2814):
    """p = (1 - eps) uniform + eps on 1000 random orbits: the output law at both
    precisions against the exact q, one BA step at sampled inputs (support and
    random) against D_0 from the exact q, and the certificate after the FP64
    step against the oracle."""
    e, n, k = e2814, 28, 14
    eps = 0.5
    rng = np.random.default_rng(2814)
    orb = e.orbits()
    pick = np.sort(rng.choice(orb.size, 1000, replace=False))
    mass = rng.random(pick.size) + 0.05
    mass /= mass.sum()
    Q = np.full(orb.size, (1.0 - eps) / 2.0 ** n)
    xs_sup, w_sup = [], []
    for i, m in zip(pick, mass):
        imgs = orbit_images(int(orb[i]), n)
        Q[i] += eps * m / len(imgs)
        xs_sup += imgs
        w_sup += [eps * m / len(imgs)] * len(imgs)
    xs_sup = np.array(xs_sup, dtype=np.uint32)
    w_sup = np.array(w_sup)
    say("BDC_(28,14): exact output law of the sparse p on the host")
    q_exact = (1.0 - eps) * 2.0 ** -k + oracle.sparse_output_law(n, k, xs_sup, w_sup)
    e.set_p_orbits(Q)
    np.testing.assert_allclose(e.log_output(), np.log(q_exact), rtol=0, atol=1e-12)
    # the output-law building block on the full-space weights, both precisions
    w = np.full(1 << n, (1.0 - eps) / 2.0 ** n)
    np.add.at(w, xs_sup.astype(np.int64), w_sup)
    # split16: p's limbs hold 45 bits relative to the maximum of its 128 x 32
    # tile, and a support string (~1e-4) shares its tile with uniform entries
    # (~2e-9), so those carry ~29 bits: |dq| <= 2^-45 * tile max * sum W over
    # the ~12% of inputs in such tiles, ~1.2e-10 relative
    for prec, rtol in ((bdc.PREC_FP64, 1e-13), (bdc.PREC_SPLIT16, 5e-10)):
        say(f"BDC_(28,14): output law of the sparse p at precision {prec}")
        np.testing.assert_allclose(e.output_law(w, prec), q_exact, rtol=rtol)
    del w
    xs = np.concatenate([xs_sup[rng.choice(xs_sup.size, 6, replace=False)], random_inputs(n, 6, 5)])
    logq = np.log(q_exact)
    D0 = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
    for prec, tol in ((bdc.PREC_SPLIT16, 1e-9), (bdc.PREC_FP64, 1e-10)):
        say(f"BDC_(28,14): one step from the sparse p at precision {prec}")
        e.set_p_orbits(Q)
        lp0 = np.log(e.sample(bdc.SAMPLE_P, xs))
        st = e.iterate(1, prec)
        lp1 = np.log(e.sample(bdc.SAMPLE_P, xs))
        np.testing.assert_allclose(lp1 - lp0 + st[0, 2], D0, rtol=0, atol=tol)
        assert st[0, 1] >= D0.max() - tol
    cert = e.certify()
    logq1 = e.log_output()
    ref = np.array([oracle.kl_x(n, k, int(x), logq1) for x in xs])
    np.testing.assert_allclose(e.sample(bdc.SAMPLE_KL, xs), ref, rtol=0, atol=1e-11)
    assert cert["upper_cert"] >= ref.max() - 1e-11
    assert abs(cert["sum_p"] - 1.0) < 1e-12


def test_bdc28_14_triangle_reads_match_mirrored(bdc, e2814, monkeypatch):
    """At the metric size (half length 14, where counts above 2048 leave
    residuals in the output law's row x1 = 0): split steps that update only the
    canonical elements, the output law reading the triangle, agree bitwise
    with steps that also write every mirror (BDC_QUARTER_TRI=0), from a
    Dirichlet(1) p and across two calls."""
    e = e2814
    out = {}
    for mode in ("1", "0"):
        say(f"BDC_(28,14): split steps, BDC_QUARTER_TRI={mode}")
        monkeypatch.setenv("BDC_QUARTER_TRI", mode)
        e.set_p_dirichlet(1.0, 3)
        st = np.vstack([e.iterate(4, bdc.PREC_SPLIT16), e.iterate(2, bdc.PREC_SPLIT16)])
        out[mode] = (st, e.certify(), e.sample(bdc.SAMPLE_P, random_inputs(28, 64, 9)))
    for a, b in zip(out["1"], out["0"]):
        if isinstance(a, dict):
            assert a == b
        else:
            assert np.array_equal(a, b)


def test_bdc28_14_from_cache_files(bdc, e2814, tmp_path):
    """Caches loaded from cache_{n'}_{k'} files (load_gpu_cache, BAA.cu:27-93;
    written by bdc_write_caches, byte-identical to the reference generator)
    give bitwise the same steps and certificate as caches generated on the
    device."""
    say("BDC_(28,14): write cache files n' <= 14")
    bdc.write_caches(14, str(tmp_path))
    say("BDC_(28,14): create from cache files")
    with bdc.Engine(28, 14, str(tmp_path)) as ef:
        for prec in (bdc.PREC_FP64, bdc.PREC_SPLIT16):
            e2814.set_p_uniform()
            a = e2814.iterate(4, prec)
            ca = e2814.certify()
            ef.set_p_uniform()
            b = ef.iterate(4, prec)
            cb = ef.certify()
            assert np.array_equal(a, b), prec
            assert ca == cb, prec


# ----------------------------------------------------------------- BDC_(28,18)
def test_bdc28_18_dual(bdc, oracle):
    """The differentiation config: theta0 = ln q after 200 split-FP16 BA steps
    from uniform; the adjoint dU/dtheta against central differences (h = 1e-6,
    64 coordinates, seed 2) and the split-FP16 gradient against FP64 for
    beta = 64 and 1024; 100 Adam steps (beta 4096, lr 1e-3) lower U and the certified
    bound, whose D(x||softmax theta) is checked against the oracle at sampled
    inputs."""
    n, k = 28, 18
    C = math.comb(n, k)
    say("BDC_(28,18): create")
    with bdc.Engine(n, k) as e:
        xs = random_inputs(n, 5, 18)
        H = np.array([oracle.entropy_x(n, k, int(x)) for x in xs])
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_ENTROPY, xs), H / C - math.log(C), rtol=0, atol=1e-12)
        say("BDC_(28,18): 200 split-FP16 BA steps")
        e.set_p_uniform()
        e.iterate(200, bdc.PREC_SPLIT16)
        theta0 = e.log_output()
        ub0 = e.certify()["upper_cert"]
        rng = np.random.default_rng(2)
        coords = rng.choice(1 << k, 64, replace=False)
        h = 1e-6
        for beta in (64.0, 1024.0):
            say(f"BDC_(28,18): gradient checks at beta {beta}")
            u64, g64 = e.dual(theta0, beta, bdc.PREC_FP64)
            u16, g16 = e.dual(theta0, beta, bdc.PREC_SPLIT16)
            assert abs(u16 - u64) < 1e-10
            assert np.abs(g16 - g64).max() < 1e-12
            fd = []
            for y in coords:
                tp, tm = theta0.copy(), theta0.copy()
                tp[y] += h
                tm[y] -= h
                fd.append((e.dual(tp, beta, bdc.PREC_FP64, want_grad=False) -
                           e.dual(tm, beta, bdc.PREC_FP64, want_grad=False)) / (2 * h))
            # central differences carry O(h^2) truncation and ~1e-16/h rounding
            assert np.abs(np.array(fd) - g64[coords]).max() < 2e-8, (beta, np.abs(np.array(fd) - g64[coords]).max())
            # U_beta >= max_x D >= I(p): the smoothed bound sits above the certificate
            assert u64 >= ub0 - 1e-12
        say("BDC_(28,18): 100 Adam steps at beta 4096")
        th, uh = e.refine_dual(theta0, 4096.0, 100, 1e-3, bdc.PREC_SPLIT16)
        assert uh[-1] < uh[0]
        ub = e.certify_q(th)
        assert ub < ub0
        logq = th - th.max()
        logq -= math.log(np.exp(logq).sum())
        ref = np.array([oracle.kl_x(n, k, int(x), logq) for x in xs])
        np.testing.assert_allclose(e.sample(bdc.SAMPLE_KL, xs), ref, rtol=0, atol=1e-11)
        assert ub >= ref.max() - 1e-11
