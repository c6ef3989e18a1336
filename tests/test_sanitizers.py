"""AddressSanitizer and UndefinedBehaviorSanitizer runs of the host code
(no GPU): the FP64 oracle (every entry point on small odd-sized problems) and
the library's host-only C-ABI sources (.bin reader / writer, normalisation
median), each built with gcc / g++ -fsanitize=address,undefined from its own
sources into a temporary directory and run; a sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _build_run(cmd, exe, args, tmp_path):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("the compiler has no sanitizer runtime: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)] + args, capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    o = os.path.join(ROOT, "oracle")
    exe = tmp_path / "oracle_check"
    cmd = ["gcc", "-std=gnu11"] + SAN + [os.path.join(ROOT, "tests", "sanitize", "oracle_check.c"),
                                         os.path.join(o, "avionic_oracle.c"),
                                         os.path.join(o, "oracle_nufft.c"),
                                         os.path.join(o, "oracle_coil.c"), "-o", str(exe), "-lm"]
    assert "oracle_check: ok" in _build_run(cmd, exe, [], tmp_path)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_capi_under_asan_ubsan(tmp_path):
    c = os.path.join(ROOT, "avionic_amd", "csrc")
    exe = tmp_path / "host_check"
    cmd = ["g++", "-std=c++17"] + SAN + [os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
                                         os.path.join(c, "binio.cpp"), os.path.join(c, "normalize.cpp"),
                                         "-o", str(exe)]
    assert "host_check: ok" in _build_run(cmd, exe, [str(tmp_path)], tmp_path)
