"""The host-side catalog reader and .npy writer (csrc/catalog.cpp) built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and
driven through ordinary, malformed, empty, binary and oversized inputs and
bad arguments (tests/sanitize/catalog_check.cpp): every call returns a
status, and the sanitizers report nothing."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_catalog_io_under_address_and_ub_sanitizers(tmp_path):
    exe = tmp_path / "catalog_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-pthread", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "baorec.jl_amd", "csrc", "catalog.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "catalog_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([str(exe), str(tmp_path)], env=env, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
