"""The C-ABI library loads and exports every entry point declared in include/sdpvibe_hip.h (no GPU
needed: only symbol lookup and the version call)."""
import ctypes
import os
import re

import pytest

from sdpvibe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "sdpvibe_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|uint64_t|void\s*\*)\s*(sdpv_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "sdpv_rns_gemm" in names and "sdpv_rns_crt" in names and len(names) >= 14


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsdpvibe_hip.so is not built (run python -m sdpvibe_amd.build)")
    h = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(h, name), name
    assert set(_declared()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"
    assert _lib.lib().sdpv_version() == _lib.EXPECTED_VERSION


def test_bad_arguments_are_rejected_without_launch():
    lib = _lib.lib()
    # negative sizes / bad modes return SDPV_EARG before touching the device
    assert lib.sdpv_shift_round(None, 1, -1, None, None, 1, None, 0, None) == -1
    assert lib.sdpv_add_same(None, None, None, 1, 4, 3, None) == -1
    assert lib.sdpv_rns_gemm(None, None, None, None, 1, 1, 4, 4, 100, None, 0, None, None, None, 0, 0, -1, None) == -2
    assert lib.sdpv_rns_gemm(None, None, None, None, 1, 1, 4, 4, 64, None, 0, None, None, None, 0, 0, 14, None) == -1
    # the syrk mode (2) takes square tiles only: configuration 12 is 64 x 128
    assert lib.sdpv_rns_gemm(None, None, None, None, 1, 1, 4, 4, 64, None, 0, None, None, None, 0, 2, 12, None) == -1
    assert lib.sdpv_mpf_mul(None, None, None, None, None, None, 200, 4, 0, None) == -2
