"""Orbit representatives and sizes (backend/BAA.cu:195-258, main_BAA.cu:46-50)."""
import numpy as np
import pytest


def burnside(n):
    return (2 ** n + 2 ** ((n + 1) // 2) + (2 ** (n // 2) if n % 2 == 0 else 0)) // 4


def brute_orbit(x, n):
    s = format(x, f"0{n}b")
    c = "".join("1" if b == "0" else "0" for b in s)
    return {int(s, 2), int(s[::-1], 2), int(c, 2), int(c[::-1], 2)}


@pytest.mark.parametrize("n", list(range(1, 17)))
def test_orbit_counts(oracle, n):
    assert len(oracle.orbit_list(n)) == burnside(n)


def test_published_orbit_counts():
    assert burnside(16) == 16512 and burnside(24) == 4196352
    assert burnside(28) == 67117056 and burnside(32) == 1073774592


@pytest.mark.parametrize("n", [1, 2, 5, 8, 11])
def test_orbits_brute_force(oracle, n):
    reps = oracle.orbit_list(n)
    assert (np.diff(reps.astype(np.int64)) > 0).all()
    seen = set()
    for x in range(1 << n):
        orb = brute_orbit(x, n)
        assert oracle.rep(x, n) == min(orb)
        assert oracle.orbit_size(x, n) == len(orb)
        seen.add(min(orb))
    assert sorted(seen) == list(reps)
