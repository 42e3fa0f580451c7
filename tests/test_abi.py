"""The C-ABI library loads on a machine without a GPU and exports every
symbol include/clat.h declares.  No compute call is made here."""
import ctypes
import os
import re

import clat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "clat.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clat_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_binding():
    assert _declared() == sorted(clat.SYMBOLS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(clat.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_strings():
    L = clat.lib()
    assert L.clat_version().decode().startswith("clat")
    assert isinstance(L.clat_last_error(), bytes)
    assert clat.kernel_kinds()[:5] == ["xfwd", "yfwd", "zconv", "yinv", "xinv"]


def test_null_handles_are_rejected():
    L = clat.lib()
    assert L.clat_prepare(None) == 1
    assert b"invalid" in L.clat_last_error()
    assert L.clat_destroy(None) == 0
