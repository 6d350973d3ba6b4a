"""Cache files handed to bdc_create (cache_io.h:7-44 format, loaded by
load_gpu_cache in backend/BAA.cu:27-93) are validated before any device work:
a missing, truncated or oversized file, or a count the u16 device layout cannot
hold, returns BDC_ERR_IO with a message.  These run on the CPU: the checks come
before the context touches HIP, so a valid directory gets as far as the device
lookup (BDC_ERR_HIP without a GPU, a context with one).
"""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def bdc():
    import bdc_gpu
    bdc_gpu.lib()
    return bdc_gpu


def write_caches(oracle, d, nmax):
    """cache_{n',k'} for 0 <= k' < n' <= nmax written from the oracle's counts
    (byte-identical to the reference generator's files, checked in
    test_oracle_counts.py)"""
    for n in range(1, nmax + 1):
        for k in range(n):
            oracle.cache(n, k).astype("<u4").tofile(str(d / f"cache_{n}_{k}"))


def create(bdc, n, k, d):
    ctx = ctypes.c_void_p()
    rc = bdc.lib().bdc_create(n, k, str(d).encode(), ctypes.byref(ctx))
    msg = bdc.lib().bdc_last_error().decode()
    if rc == bdc.OK:
        bdc.lib().bdc_destroy(ctx)
    return rc, msg


def test_valid_directory_passes_validation(bdc, oracle, tmp_path):
    write_caches(oracle, tmp_path, 8)
    rc, msg = create(bdc, 16, 8, tmp_path)
    # without a GPU the context stops at the device lookup, after the files passed
    assert rc in (bdc.OK, bdc.ERR_HIP), (rc, msg)


@pytest.mark.parametrize("case", ["missing", "truncated", "oversized", "count_too_large"])
def test_bad_cache_file(bdc, oracle, tmp_path, case):
    write_caches(oracle, tmp_path, 8)
    f = tmp_path / "cache_8_5"
    data = f.read_bytes()
    if case == "missing":
        f.unlink()
    elif case == "truncated":
        f.write_bytes(data[:-1])
    elif case == "oversized":
        f.write_bytes(data + b"\0\0\0\0")
    else:
        a = np.frombuffer(data, dtype="<u4").copy()
        a[1234] = 70000
        a.tofile(str(f))
    rc, msg = create(bdc, 16, 8, tmp_path)
    assert rc == bdc.ERR_IO, (rc, msg)
    assert "cache_8_5" in msg, msg
    if case == "count_too_large":
        assert "70000" in msg and "1234" in msg, msg
    if case in ("truncated", "oversized"):
        assert "expected 8192 uint32 counts" in msg, msg


def test_only_needed_files_are_read(bdc, oracle, tmp_path):
    """BDC_(10,3) reads cache_5_j for j <= 3 only (the split sum of
    backend/BAA.cu:355-392 never reaches longer outputs): a broken cache_5_4
    does not matter, a broken cache_5_3 does."""
    write_caches(oracle, tmp_path, 5)
    (tmp_path / "cache_5_4").write_bytes(b"\1")
    rc, msg = create(bdc, 10, 3, tmp_path)
    assert rc in (bdc.OK, bdc.ERR_HIP), (rc, msg)
    (tmp_path / "cache_5_3").write_bytes(b"\1")
    rc, msg = create(bdc, 10, 3, tmp_path)
    assert rc == bdc.ERR_IO and "cache_5_3" in msg, (rc, msg)
