"""CPU check of the output law's triangle reads (k_split_g1 with TRI, the
quarter-layout BA step): for every block, thread and stage, the byte offsets the
kernel forms -- in place, per element in place or from the partner, or from the
partner as 16-byte pairs -- evaluated with the kernel's 32-bit wrap-around, must
stay inside the half space and land on the canonical representative of the
element the thread multiplies.  The stage plan (which of the kernel's three loops
loads each stage, four stages ahead) follows the kernel's loop bounds.  Offsets right of the middle column are sums of
a per-lane part that can be negative: they wrap in 32 bits before they meet a
pointer."""
import numpy as np
import pytest

BK = 64


def g1_splits(rows, blocks):
    ks = 1
    while ks < 8 and blocks * ks < 512 and (rows // 128) % (ks * 2) == 0 and rows // (ks * 2) >= 1024:
        ks *= 2
    return ks


def stage_plan(nst, kbase, cmin, cmax):
    """{stage: set of load kinds}: 'I' in place, 'G' per-element select, 'T' partner pairs."""
    sin = (cmin - (BK - 1) - kbase) // BK + 1 if cmin - (BK - 1) - kbase >= 0 else 0
    st = (cmax - kbase) // BK + 1 if cmax - kbase >= 0 else 0
    plan = {}

    def load(s, kind):
        plan.setdefault(min(s + 4, nst - 1), set()).add(kind)

    for t in range(min(4, nst)):  # prologue
        plan.setdefault(t, set()).add("G")
    s0 = 0
    for kind, more in (("I", lambda s: s + 10 <= sin), ("G", lambda s: s + 4 < st), ("T", lambda s: True)):
        while s0 + 6 <= nst and more(s0):
            for j in range(6):
                load(s0 + j, kind)
            s0 += 6
    for j in range(4):  # the last nst mod 6 stages
        if s0 + j < nst:
            load(s0 + j, "G")
    return plan


def check(n):
    N = 1 << (n // 2)
    M = N // 2
    ncols, kdim = N, M
    ks = g1_splits(M, ncols // 32)
    r = np.arange(M)[:, None]
    c = np.arange(N)[None, :]
    left_m = c < M
    canon = np.where(left_m, r <= c, r + c <= N - 1)
    ident = r * ncols + c
    phys = np.where(canon, ident, -1).reshape(-1)  # only canonical elements are current
    want = np.where(canon, ident, np.where(left_m, c * ncols + r, (N - 1 - c) * ncols + (N - 1 - r)))
    nbytes = M * ncols * 8
    tid = np.arange(512)
    wave, lane = tid >> 6, tid & 63
    rp = 8 * (wave >> 1) + (lane & 7)
    cp = 8 * (wave & 1) + (lane >> 3)
    wrap = np.uint64(1 << 32)
    seen = set()
    for tile in range(ncols // 32):
        n0 = tile * 32
        left = n0 < M
        offx = (2 * rp * ncols + 2 * cp) * 8
        offt = (2 * cp * ncols + 2 * rp) * 8 if left else ((30 - 2 * cp) * ncols + 30 - 2 * rp) * 8
        offt = offt.astype(np.int64) % (1 << 32)  # uint32 in the kernel
        cmin, cmax = (n0, n0 + 31) if left else (ncols - 32 - n0, ncols - 1 - n0)
        for kz in range(ks):
            krows = kdim // ks
            kbase, nst = kz * krows, krows // BK
            for s, kinds in stage_plan(nst, kbase, cmin, cmax).items():
                r0 = kbase + s * BK
                for kind in kinds:
                    seen.add(kind)

                    def ut(i, q):
                        if left:
                            return ((n0 + q) * ncols + r0 + i) * 8
                        return ((ncols - 31 - n0 - q) * ncols + ncols - 31 - r0 - i) * 8

                    for i in range(2):
                        for q in range(2):
                            if kind == "I":
                                a = (r0 * ncols + n0) * 8 + i * ncols * 8 + offx + 8 * q
                            elif kind == "T":  # rows (0, 1) left, (1, 0) right, at ascending addresses
                                pair = (ut(0 if left else 1, q) + offt) % (1 << 32)
                                assert np.all(pair % 16 == 0)
                                a = pair + 8 * (i if left else 1 - i)
                            else:
                                on = ((r0 + i) * ncols + n0 + q) * 8 + offx
                                ot = (ut(i, q) + offt) % (1 << 32)
                                dq = 2 * rp - ((n0 + 2 * cp + q) if left else (ncols - 1 - n0 - 2 * cp - q))
                                a = np.where(dq <= -(r0 + i), on, ot)
                            a = np.asarray(a, dtype=np.int64)
                            assert np.all((a >= 0) & (a < nbytes) & (a % 8 == 0)), (n, kind, tile, kz, s)
                            R, C = r0 + 2 * rp + i, n0 + 2 * cp + q
                            assert np.array_equal(phys[a // 8], want[R, C]), (n, kind, tile, kz, s, i, q)
    assert np.uint64(nbytes) < wrap
    return ks, seen


@pytest.mark.parametrize("n,kinds", [(16, {"G"}), (20, {"G", "T"}), (24, {"G", "I", "T"})])
def test_triangle_reads_land_on_canonical_partners(n, kinds):
    ks, seen = check(n)
    assert seen == kinds
    if n == 24:
        assert ks == 2  # a second K slab (kbase > 0) is covered
