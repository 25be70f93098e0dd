"""The file-format code built with AddressSanitizer and UBSan and driven by
tests/sanitize/formats_check.cpp: round trips, every truncation, random
corruption and random headers end in a status code without any memory or
undefined-behaviour error; the result document is exact in length and never
holds a non-JSON number.  Host code only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_formats_under_sanitizers(tmp_path):
    exe = str(tmp_path / "formats_check")
    # the sanitizer runtimes are linked statically, so they initialise first in the process
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-static-libasan", "-static-libubsan",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "dft_amd", "host"),
           os.path.join(ROOT, "tests", "sanitize", "formats_check.cpp"),
           os.path.join(ROOT, "dft_amd", "host", "formats.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0 and run.stdout.strip() == "ok", run.stdout + run.stderr
