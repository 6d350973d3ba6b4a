"""The oracle's subsequence counts and cache files against the reference
generator (golden digests and raw files), brute force and closed forms."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import dense_ba


def test_count_matches_brute_force(oracle):
    rng = np.random.default_rng(0)
    for n in range(0, 11):
        for k in range(0, n + 1):
            for _ in range(6):
                x = int(rng.integers(0, 1 << n)) if n else 0
                y = int(rng.integers(0, 1 << k)) if k else 0
                assert oracle.count(x, n, y, k) == dense_ba.brute_count(x, n, y, k)


def test_count_matrix_helper_matches_brute_force():
    n, k = 7, 3
    C = dense_ba.count_matrix(n, k)
    for x in range(1 << n):
        for y in range(1 << k):
            assert C[x, y] == dense_ba.brute_count(x, n, y, k)


def test_cache_raw_files_byte_identical(oracle, golden_dir):
    raw = os.path.join(golden_dir, "raw")
    files = sorted(os.listdir(raw))
    assert len(files) == sum(n for n in range(1, 7))
    for name in files:
        _, n, k = name.split("_")
        n, k = int(n), int(k)
        with open(os.path.join(raw, name), "rb") as f:
            data = f.read()
        assert oracle.cache(n, k).tobytes() == data, name


def test_cache_digests_match_reference_generator(oracle, golden_dir):
    with open(os.path.join(golden_dir, "cache_sha256.json")) as f:
        ref = json.load(f)["sha256"]
    assert len(ref) == sum(n for n in range(1, 15))
    for key, digest in ref.items():
        n, k = map(int, key.split("_"))
        if n > 12:
            continue  # n' = 13, 14 are checked on the GPU path and in test_cache_digests_large
        assert hashlib.sha256(oracle.cache(n, k).tobytes()).hexdigest() == digest, key


@pytest.mark.parametrize("n,k", [(13, 12), (14, 7), (14, 13)])
def test_cache_digests_large(oracle, golden_dir, n, k):
    with open(os.path.join(golden_dir, "cache_sha256.json")) as f:
        ref = json.load(f)["sha256"]
    assert hashlib.sha256(oracle.cache(n, k).tobytes()).hexdigest() == ref[f"{n}_{k}"]


@pytest.mark.parametrize("n,k", [(8, 3), (9, 5), (10, 9)])
def test_cache_invariants(oracle, n, k):
    c = oracle.cache(n, k).reshape(1 << n, 1 << k).astype(np.int64)
    assert (c.sum(axis=1) == math.comb(n, k)).all()  # row sums
    assert (c.sum(axis=0) == math.comb(n, k) * 2 ** (n - k)).all()  # column sums
    assert c[0, 0] == math.comb(n, k)
    # complement of both strings, reversal of both strings
    xs, ys = np.arange(1 << n), np.arange(1 << k)
    comp_x, comp_y = xs ^ ((1 << n) - 1), ys ^ ((1 << k) - 1)
    assert (c[np.ix_(comp_x, comp_y)] == c).all()
    rev = lambda v, m: int(format(v, f"0{m}b")[::-1], 2) if m else 0
    rx = np.array([rev(v, n) for v in xs])
    ry = np.array([rev(v, k) for v in ys])
    assert (c[np.ix_(rx, ry)] == c).all()


def test_cache_k0_all_ones(oracle):
    for n in range(1, 9):
        assert (oracle.cache(n, 0) == 1).all()


def test_entropy_x_matches_dense_channel(oracle):
    # sum_y W ln W over the dense channel equals H(x)/C - ln C
    for n, k in [(10, 5), (9, 3), (8, 7)]:
        W = dense_ba.channel(n, k)
        C = math.comb(n, k)
        with np.errstate(divide="ignore", invalid="ignore"):
            ref = np.where(W > 0, W * np.log(np.where(W > 0, W, 1.0)), 0.0).sum(axis=1)
        got = np.array([oracle.entropy_x(n, k, x) / C - math.log(C) for x in range(1 << n)])
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


def test_entropy_x_closed_form_runs(oracle):
    # x = 0^a 1^b: the length-k subsequences are 0^i 1^(k-i), each C(a,i) C(b,k-i)
    # times, so H = sum_i c_i ln c_i; checked at n = 32 (a = b = 16, k = 16)
    for a, b, k in [(16, 16, 16), (15, 14, 14), (3, 5, 4)]:
        n = a + b
        x = (1 << b) - 1
        ref = 0.0
        for i in range(k + 1):
            c = math.comb(a, i) * math.comb(b, k - i)
            if c > 1:
                ref += c * math.log(c)
        assert abs(oracle.entropy_x(n, k, x) - ref) <= 1e-12 * abs(ref)


def test_kl_x_matches_full_kl(oracle):
    n, k = 10, 4
    rng = np.random.default_rng(3)
    q = rng.random(1 << k) + 0.1
    logq = np.log(q / q.sum())
    D = oracle.full_kl(n, k, logq)
    for x in [0, 1, 77, 512, 1023]:
        assert abs(oracle.kl_x(n, k, x, logq) - D[x]) <= 1e-13
    # uniform q: D(x||q) = H(x)/C - ln C + k ln 2
    C = math.comb(n, k)
    u = np.full(1 << k, -k * math.log(2.0))
    for x in [5, 300]:
        ref = oracle.entropy_x(n, k, x) / C - math.log(C) + k * math.log(2.0)
        assert abs(oracle.kl_x(n, k, x, u) - ref) <= 1e-13


def test_single_input_sums_for_k_near_n(oracle):
    """entropy_x / kl_x visit the distinct subsequences (the reference's
    enumeration, BAA.cu:261-335): against a brute-force sweep over every output
    for k close to n and k = n, and the single-deletion closed form at n = 29:
    the distinct y of length n - 1 are one per run of x, reached ell times for a
    run of length ell, so H(x) = sum over runs of ell ln ell."""
    rng = np.random.default_rng(11)
    for n, k in [(13, 11), (12, 12), (11, 10)]:
        logq = np.log(rng.dirichlet(np.ones(1 << k)))
        C = math.comb(n, k)
        for x in [0, (1 << n) - 1, int(rng.integers(1 << n)), int(rng.integers(1 << n))]:
            cs = np.array([oracle.count(x, n, y, k) for y in range(1 << k)], dtype=np.float64)
            nz = cs > 0
            H = float(np.sum(cs[cs > 1] * np.log(cs[cs > 1])))
            D = float(np.sum(cs[nz] / C * (np.log(cs[nz]) - math.log(C) - logq[nz])))
            assert abs(oracle.entropy_x(n, k, x) - H) <= 1e-12 * max(1.0, H)
            assert abs(oracle.kl_x(n, k, x, logq) - D) <= 1e-12
    n = 29
    for x in [0x0ABCDEF1, 0x155AA55A, 0, 0x1FFFFFFF, 0x0F0F0F0F]:
        bits = format(x, f"0{n}b")
        runs, cur = [], 1
        for i in range(1, n):
            if bits[i] == bits[i - 1]:
                cur += 1
            else:
                runs.append(cur)
                cur = 1
        runs.append(cur)
        ref = sum(l * math.log(l) for l in runs if l > 1)
        assert abs(oracle.entropy_x(n, n - 1, x) - ref) <= 1e-12 * max(1.0, ref)


def test_sparse_output_law(oracle):
    """q of a weight vector on a few inputs against the full-space sum"""
    rng = np.random.default_rng(5)
    for n, k in [(10, 5), (12, 7), (11, 11)]:
        xs = rng.choice(1 << n, 40, replace=False).astype(np.uint32)
        w = rng.random(40)
        full = np.zeros(1 << n)
        full[xs] = w
        np.testing.assert_allclose(oracle.sparse_output_law(n, k, xs, w), oracle.full_output_law(n, k, full),
                                   rtol=1e-13, atol=1e-300)
