"""The C-ABI library loads and exports every symbol include/bdc.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(REPO, "include", "bdc.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bdc_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_python_binding():
    import bdc_gpu
    assert declared_symbols() == sorted(bdc_gpu.EXPORTS)


def test_library_exports_every_symbol():
    import bdc_gpu
    if not os.path.exists(bdc_gpu.LIB_PATH):
        bdc_gpu.build()
    L = ctypes.CDLL(bdc_gpu.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_supported_validation_without_gpu():
    import bdc_gpu
    L = bdc_gpu.lib()
    assert L.bdc_supported(16, 8) == 1
    assert L.bdc_supported(16, 16) == 1  # k = n (the identity channel, C = n bits)
    assert L.bdc_supported(16, 0) == 0
    assert L.bdc_supported(16, 17) == 0
    assert b"k must be in [1, n]" in L.bdc_last_error()
    # inputs up to 2^32 (the scale config BDC_(32,16)), outputs up to 2^30
    assert L.bdc_supported(32, 16) == 1
    assert L.bdc_supported(29, 14) == 1
    assert L.bdc_supported(29, 28) == 1
    assert L.bdc_supported(31, 30) == 1
    assert L.bdc_supported(33, 16) == 0
    assert b"[2, 32]" in L.bdc_last_error()
    assert L.bdc_supported(32, 31) == 0
    assert b"k <= 30" in L.bdc_last_error()
    assert L.bdc_version() >= 1


def test_argument_errors_without_gpu():
    """Every entry point checks its arguments before any GPU call: a status
    code and a message, never an exit (the reference exits, BAA.cuh:5-13)."""
    import bdc_gpu
    L = bdc_gpu.lib()
    ctx = ctypes.c_void_p(1234)
    assert L.bdc_create(33, 16, None, ctypes.byref(ctx)) == bdc_gpu.ERR_ARG
    assert ctx.value is None  # cleared on failure
    assert b"n must be in [2, 32]" in L.bdc_last_error()
    assert L.bdc_create(28, 29, None, ctypes.byref(ctx)) == bdc_gpu.ERR_ARG
    assert b"k must be in [1, n]" in L.bdc_last_error()
    assert L.bdc_create(28, 14, None, None) == bdc_gpu.ERR_ARG
    assert L.bdc_iterate(None, bdc_gpu.PREC_FP64, 1, 0.0, None, None) == bdc_gpu.ERR_ARG
    assert L.bdc_certify(None, None) == bdc_gpu.ERR_ARG
    assert L.bdc_set_p_uniform(None) == bdc_gpu.ERR_ARG
    assert L.bdc_set_p_dirichlet(None, 1.0, 0) == bdc_gpu.ERR_ARG
    u = ctypes.c_double()
    theta = (ctypes.c_double * 4)()
    assert L.bdc_dual(None, bdc_gpu.PREC_FP64, theta, 64.0, ctypes.byref(u), None) == bdc_gpu.ERR_ARG
    assert L.bdc_refine_dual(None, bdc_gpu.PREC_FP64, theta, 64.0, 1, 1e-3, None) == bdc_gpu.ERR_ARG
    assert L.bdc_destroy(None) == bdc_gpu.OK
