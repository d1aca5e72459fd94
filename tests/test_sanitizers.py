"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the oracle
(oracle/hdg_oracle.c) through its whole pipeline, and libhdg's host-only
code (the mesh generator, the reference and Schur-form tables, the serving
arena's free list), each built with gcc/g++ and the sanitizers and run by a
driver under tests/sanitize/.  Device code is not built here: GPU
sanitizers are not available for gfx950 in this setting."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

SAN = ["-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _build_and_run(cmd, exe):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr and ("cannot find" in r.stderr or "unrecognized" in r.stderr):
        pytest.skip("compiler without sanitizer runtimes")
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_san")
    cmd = ["gcc", "-std=gnu11", *SAN, "-fopenmp", "-I", os.path.join(REPO, "oracle"),
           os.path.join(REPO, "tests", "sanitize", "oracle_driver.c"), os.path.join(REPO, "oracle", "hdg_oracle.c"),
           "-lm", "-o", exe]
    assert "ok" in _build_and_run(cmd, exe)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_code_under_asan_ubsan(tmp_path):
    csrc = os.path.join(REPO, "hdiscontinuousgalerkin.jl_amd", "csrc")
    exe = str(tmp_path / "host_san")
    cmd = ["g++", "-std=c++17", *SAN, "-I", os.path.join(REPO, "include"), "-I", csrc,
           os.path.join(REPO, "tests", "sanitize", "host_driver.cpp"), os.path.join(csrc, "hdg_mesh.cpp"),
           os.path.join(csrc, "hdg_tables.cpp"), "-o", exe]
    assert "ok" in _build_and_run(cmd, exe)
