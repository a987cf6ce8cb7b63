"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, no GPU work).

* The CPU oracle (oracle/oracle.cpp) with oracle/sanitize_main.cpp, which walks every major entry point on small
  grids.
* The reference-named C++ API (classicaldft_amd/host/classicaldft.cpp) with tests/sanitize_host.cpp, which runs its
  host-side parts (vectors, lattice, density files, species, potentials) against the uninstrumented libcdft.so.

Each driver is compiled with g++ -fsanitize=address,undefined into a temporary directory and must exit 0 with no
sanitizer report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)


def _check(res):
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_san")
    _check(_run(["g++", *SAN, "-fopenmp", "-std=gnu++17", "-fext-numeric-literals", "-ffp-contract=off", "-o", exe,
                 os.path.join(ROOT, "oracle", "sanitize_main.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
                 "-lquadmath"]))
    out = _check(_run([exe], env=ENV))
    assert "sanitize: ok" in out


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_api_under_asan_and_ubsan(tmp_path):
    lib = os.path.join(ROOT, "classicaldft_amd")
    if not os.path.exists(os.path.join(lib, "libcdft.so")):
        pytest.skip("libcdft.so is not built")
    exe = str(tmp_path / "host_san")
    _check(_run(["g++", *SAN, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe,
                 os.path.join(ROOT, "tests", "sanitize_host.cpp"), os.path.join(lib, "host", "classicaldft.cpp"),
                 "-L", lib, "-lcdft", "-Wl,-rpath," + lib]))
    out = _check(_run([exe, str(tmp_path)], env=ENV))
    assert "sanitize host: ok" in out
