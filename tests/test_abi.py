"""The C-ABI libraries load and export every symbol their headers declare (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdft_[a-z0-9_]+)\s*\(", text)))


def test_libcdft_exports_header():
    lib = ctypes.CDLL(os.path.join(ROOT, "classicaldft_amd", "libcdft.so"))
    names = declared("cdft.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_import():
    import classicaldft_amd as c
    assert os.path.exists(c.library_path())
    assert c.WHITEBEAR_I == 3 and c.FP64 == 0 and c.MIXED == 1
