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
    assert L.bdc_supported(16, 16) == 0
    assert L.bdc_supported(16, 0) == 0
    # inputs up to 2^32 (the scale config BDC_(32,16)), outputs up to 2^20
    assert L.bdc_supported(32, 16) == 1
    assert L.bdc_supported(29, 14) == 1
    assert L.bdc_supported(33, 16) == 0
    assert b"[2, 32]" in L.bdc_last_error()
    assert L.bdc_supported(32, 21) == 0
    assert L.bdc_version() >= 1
