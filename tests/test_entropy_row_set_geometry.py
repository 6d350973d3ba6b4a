"""CPU check of the entropy kernel's grouped x2 row-set build (k_entropy_reps,
build_b_grouped over the tables stage_b_tables copies into LDS): for the full
layout (r = 0) and the quadrant layout (r = 1, entries restricted to z's
bottom bit), the index arithmetic of every group of 8 entries is evaluated as
the kernel forms it, and must

- write every entry of the levels it builds exactly once, 16-byte aligned;
- read the low-half table 16-byte aligned and inside the staged entries, and
  the high-half table inside its (2 << a) - 1 entries;
- give c(x2, z), the number of occurrences of z as a subsequence of x2,
  computed here by brute force (a DP over x2's symbols for every z at once),
  with every packed u16 partial sum below 2^16.
"""
import numpy as np
import pytest


def lev_off(l, r):
    return (1 << (l - 1) if l else 0) if r else 1 << l


def counts_all(s, L, m):
    """c(s, z) for every z of length m (MSB first), s of length L"""
    z = np.arange(1 << m, dtype=np.int64)
    dp = np.zeros((m + 1, 1 << m), dtype=np.int64)
    dp[0] = 1
    for i in range(L):
        si = (s >> (L - 1 - i)) & 1
        for j in range(min(i + 1, m), 0, -1):
            dp[j] += dp[j - 1] * (((z >> (m - j)) & 1) == si)
    return dp[m]


def half_table(h, l):
    """T[2^d - 1 + w] = c(h, w) for |w| = d <= l (the kernel's half tables)"""
    return np.concatenate([counts_all(h, l, d) for d in range(l + 1)])


def check(n2, x2, r, bit, lo, hi):
    a, bl = (n2 + 1) // 2, n2 // 2
    D0, M0 = (4, 4) if r else (3, 3)
    assert lo >= M0
    thb = half_table(x2 >> bl, a)  # the low u16 half of the staged u32
    tl = half_table(x2 & ((1 << bl) - 1), bl)
    nr = 1 << bl if r else 2 << bl
    tlr = np.zeros(nr, dtype=np.int64)
    for f in range(nr):
        if not r and not f:
            continue
        d = (f.bit_length() if f else 0) if r else f.bit_length() - 1
        e = f - lev_off(d, r)
        w = ((e << 1) | bit) if (r and d) else e
        tlr[f] = tl[(1 << d) - 1 + w]
    f0, f1 = lev_off(lo, r), lev_off(hi, r) + ((1 << (hi - 1)) if r else 1 << hi)
    written = np.zeros(f1, dtype=np.int64)
    B = np.zeros(f1, dtype=np.int64)
    for f in range(f0, f1, 8):
        assert f % 8 == 0  # 16-byte store
        m = f.bit_length() if r else f.bit_length() - 1
        e0 = f - lev_off(m, r)
        z0 = ((e0 << 1) | bit) if r else e0
        acc = np.zeros(8, dtype=np.int64)
        for i in range(max(0, m - bl), min(m, a) + 1):
            d = m - i
            if d >= D0:
                ih = (1 << i) - 1 + (z0 >> d)
                assert ih < (2 << a) - 1
                il = lev_off(d, r) + ((z0 & ((1 << d) - 1)) >> r)
                assert il % 8 == 0 and il + 8 <= nr  # one 16-byte LDS load
                acc += thb[ih] * tlr[il:il + 8]
            else:
                for j in range(8):
                    z = z0 + (j << r)
                    ih, il = (1 << i) - 1 + (z >> d), (1 << d) - 1 + (z & ((1 << d) - 1))
                    assert ih < (2 << a) - 1 and il < (2 << bl) - 1
                    acc[j] += thb[ih] * tl[il]
            assert acc.max() < 1 << 16  # packed u16 sums stay exact
        B[f:f + 8] = acc
        written[f:f + 8] += 1
    assert np.all(written[f0:f1] == 1)
    for m in range(lo, hi + 1):
        ref = counts_all(x2, n2, m)
        cnt = (1 << (m - 1)) if r else 1 << m
        z = ((np.arange(cnt) << 1) | bit) if r else np.arange(cnt)
        assert np.array_equal(B[lev_off(m, r):lev_off(m, r) + cnt], ref[z]), (n2, x2, r, bit, m)


@pytest.mark.parametrize("n2", [6, 7, 10, 13, 16])
@pytest.mark.parametrize("r", [0, 1])
def test_grouped_row_sets_match_subsequence_counts(n2, r):
    rng = np.random.default_rng(n2 * 2 + r)
    hi = n2 - 1
    lo = 4 if r else 3
    if hi < lo:
        pytest.skip("no level of 8 or more entries")
    xs = [0, (1 << n2) - 1, int("10" * (n2 // 2) + "1" * (n2 % 2), 2)] + [int(v) for v in rng.integers(0, 1 << n2, 2)]
    for x2 in xs:
        for bit in ((0, 1) if r else (0,)):
            check(n2, x2, r, bit, lo, hi)
