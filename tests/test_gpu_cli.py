"""The main_BAA command-line driver (backend/main_BAA.cu:134-170 surface) on the
GPU: iteration count and saved distribution against the oracle, checkpoint
sidecar, resume, and rejection of malformed distribution files."""
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "bdc-upper-bounds-gpu_amd", "bin", "main_BAA")
LN2 = math.log(2.0)


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def run(cwd, *args):
    if not os.path.exists(EXE):
        raise RuntimeError(f"{EXE} is not built (run __graft_entry__.build())")
    return subprocess.run([EXE, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=120)


def read_meta(path):
    out = {}
    for line in open(path):
        key, val = line.split()
        out[key] = val
    return out


def oracle_run(oracle, n, k, tol, start=None):
    ref = oracle.BA(n, k)
    if start is not None:
        ref.set_q(start)
    it = 0
    while True:
        it += 1
        if ref.iterate() < tol:
            return ref, it


def test_main_baa_run_save_resume(oracle, tmp_path):
    n, k = 12, 6
    say("main_BAA 12 6 0 1e-3 7")
    r = run(tmp_path, n, k, 0, 1e-3, 7)
    assert r.returncode == 0, r.stderr
    ref, iters = oracle_run(oracle, n, k, 1e-3)
    assert len(re.findall(r"^Iteration \d+: mx = ", r.stdout, re.M)) == iters
    dist = tmp_path / "distributions" / f"dist_{n}_{k}.txt"
    q = np.loadtxt(dist)
    assert q.shape == (ref.num_orbits,)
    np.testing.assert_allclose(q, ref.get_q(), rtol=1e-10)
    meta = read_meta(str(dist) + ".meta")
    assert int(meta["n"]) == n and int(meta["k"]) == k and meta["precision"] == "fp64"
    assert int(meta["iteration"]) == iters and float(meta["mx"]) < 1e-3
    c = float(re.search(r"Computed C_\{12,6\} = ([0-9.]+) bits", r.stdout).group(1))
    assert abs(c - ref.rate() / LN2) < 1e-6
    ub = float(re.search(r"max_x D\(x\|\|q\) = ([0-9.]+) bits", r.stdout).group(1))
    assert abs(ub - ref.max_kl() / LN2) < 1e-9

    # resume to a tighter tolerance: the count continues and the state equals
    # an uninterrupted run of the oracle from the saved file
    say("main_BAA 12 6 1 1e-4 0 (resume)")
    r2 = run(tmp_path, n, k, 1, 1e-4, 0)
    assert r2.returncode == 0, r2.stderr
    assert f"Resuming after iteration {iters}" in r2.stdout
    first = int(re.search(r"^Iteration (\d+):", r2.stdout, re.M).group(1))
    assert first == iters + 1
    ref2, more = oracle_run(oracle, n, k, 1e-4, start=q)
    meta2 = read_meta(str(dist) + ".meta")
    assert int(meta2["iteration"]) == iters + more
    # thousands of steps drive some orbits into the subnormal range (down to
    # 1e-323, a bit or two of precision), where only an absolute floor applies
    np.testing.assert_allclose(np.loadtxt(dist), ref2.get_q(), rtol=1e-9, atol=1e-280)
    ub2 = float(re.search(r"max_x D\(x\|\|q\) = ([0-9.]+) bits", r2.stdout).group(1))
    assert abs(ub2 - ref2.max_kl() / LN2) < 1e-9
    assert not os.path.exists(str(dist) + ".tmp")


def test_main_baa_split16(oracle, tmp_path):
    n, k = 16, 8
    say("main_BAA 16 8 0 1e-2 0 split16")
    r = run(tmp_path, n, k, 0, 1e-2, 0, "split16")
    assert r.returncode == 0, r.stderr
    ref, iters = oracle_run(oracle, n, k, 1e-2)
    assert len(re.findall(r"^Iteration \d+: mx = ", r.stdout, re.M)) == iters
    ub = float(re.search(r"max_x D\(x\|\|q\) = ([0-9.]+) bits", r.stdout).group(1))
    assert abs(ub - ref.max_kl() / LN2) < 2e-9  # printed to 10 decimals
    assert read_meta(str(tmp_path / "distributions" / f"dist_{n}_{k}.txt.meta"))["precision"] == "split16"


def test_main_baa_loads_transition_counts(tmp_path):
    """With the half-length cache files in ./transition_counts (the reference's
    cache directory, BAA.cu:27-93) main_BAA loads them instead of generating
    the counts, and the run is bitwise that of a run without them."""
    import bdc_gpu
    n, k = 16, 8
    say("main_BAA 16 8 with and without ./transition_counts")
    plain = tmp_path / "plain"
    cached = tmp_path / "cached"
    plain.mkdir()
    (cached / "transition_counts").mkdir(parents=True)
    bdc_gpu.write_caches(8, str(cached / "transition_counts"))
    r1 = run(plain, n, k, 0, 1e-3, 0)
    r2 = run(cached, n, k, 0, 1e-3, 0)
    assert r1.returncode == 0 and r2.returncode == 0, (r1.stderr, r2.stderr)
    assert "Cache files not found" in r1.stdout
    assert "Cache files not found" not in r2.stdout
    d1 = (plain / "distributions" / f"dist_{n}_{k}.txt").read_bytes()
    d2 = (cached / "distributions" / f"dist_{n}_{k}.txt").read_bytes()
    assert d1 == d2
    # a truncated cache file is an error, not a silent regeneration
    f = cached / "transition_counts" / "cache_8_3"
    f.write_bytes(f.read_bytes()[:-4])
    r3 = run(cached, n, k, 0, 1e-3, 0)
    assert r3.returncode != 0 and "cache_8_3" in r3.stderr


@pytest.mark.parametrize("case", ["short", "unnormalised", "negative", "garbage"])
def test_main_baa_rejects_bad_distribution(oracle, tmp_path, case):
    n, k = 10, 5
    say(f"main_BAA bad dist {case}")
    ref = oracle.BA(n, k)
    q = ref.get_q()
    (tmp_path / "distributions").mkdir()
    dist = tmp_path / "distributions" / f"dist_{n}_{k}.txt"
    if case == "short":
        q = q[:-1]
    elif case == "unnormalised":
        q = q * 2.0
    elif case == "negative":
        q = q.copy()
        q[3] = -q[3]
    text = "".join(f"{v:.17g}\n" for v in q)
    if case == "garbage":
        text = text.replace("\n", "\nx\n", 1)
    dist.write_text(text)
    r = run(tmp_path, n, k, 1, 1e-3, 0)
    assert r.returncode != 0
    assert "Error" in r.stderr
