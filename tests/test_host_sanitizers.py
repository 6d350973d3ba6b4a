"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds two small drivers with gcc/g++ -fsanitize=address,undefined:
tests/asan/dist_io_check.cpp over main_BAA's distribution-file and sidecar I/O
(round trip and malformed files), and tests/asan/oracle_check.c over every
entry point of the CPU oracle at small sizes.  A sanitizer report fails the run.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def build_and_run(cmd, exe, args, tmp_path):
    if shutil.which(cmd[0]) is None:
        pytest.skip(f"{cmd[0]} not available")
    subprocess.run(cmd + ["-o", str(exe)], check=True, capture_output=True, text=True)
    # verify_asan_link_order=0: the environment may preload a library ahead of
    # the sanitizer runtime, which is left in place
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300, env=env, cwd=tmp_path)


def test_dist_io_under_sanitizers(tmp_path):
    src = os.path.join(REPO, "tests", "asan", "dist_io_check.cpp")
    r = build_and_run(["g++", "-std=c++17", *SAN, src, "-pthread"], tmp_path / "dist_io_check", [str(tmp_path)],
                      tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_oracle_under_sanitizers(tmp_path):
    src = [os.path.join(REPO, "tests", "asan", "oracle_check.c"), os.path.join(REPO, "oracle", "bdc_oracle.c")]
    r = build_and_run(["gcc", "-std=c11", "-fopenmp", *SAN, *src, "-lm"], tmp_path / "oracle_check", [], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle checks: ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
