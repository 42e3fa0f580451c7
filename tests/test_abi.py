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


def test_header_compiles_as_c_and_links(tmp_path):
    """A C99 program (the language of a cgo or FFI stub) includes the header,
    links the library and sees the library's error path: without a GPU,
    clat_create fails with CLAT_EDEVICE and a message; with one, it succeeds."""
    import subprocess
    src = tmp_path / "use_clat.c"
    src.write_text(
        '#include <stdio.h>\n#include "clat.h"\n'
        "int main(void) {\n"
        '  printf("%s\\n", clat_version());\n'
        "  clat_engine* e = 0;\n"
        "  int rc = clat_create(32, 32, 32, CLAT_PREC_FP64, CLAT_SCHED_SHARED, &e);\n"
        '  printf("%d %s\\n", rc, rc ? clat_last_error() : "ok");\n'
        "  if (rc == CLAT_OK) return clat_destroy(e) == CLAT_OK ? 0 : 3;\n"
        "  return rc == CLAT_EDEVICE && clat_last_error()[0] ? 0 : 4;\n"
        "}\n")
    exe = tmp_path / "use_clat"
    libdir = os.path.dirname(clat.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe), "-L", libdir, "-lclat", "-Wl,-rpath," + libdir], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("clat")
