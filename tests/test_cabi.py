"""The C ABI (include/fh_api.h) against the built library, without a GPU: the library loads,
exports every function the header declares, and the ctypes binding declares nothing else."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fh_api.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    assert "fh_gemm_launch" in names and "fh_ns_fused" in names and "fh_csr_sort" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from fasthigashi import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfh_hip.so not built")
    so = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(so, n)]
    assert not missing, missing
    assert so.fh_version() == 100
    so.fh_target.restype = ctypes.c_char_p
    assert so.fh_target() == b"gfx950"


def test_binding_matches_header():
    from fasthigashi import _lib
    assert sorted(_lib.declared_symbols()) == _declared()
