"""AddressSanitizer + UndefinedBehaviorSanitizer builds of host code (CPU only).

* The oracle built with -fsanitize=address,undefined (oracle/Makefile `san`) runs the oracle test files;
  any heap/stack overflow, use-after-free or undefined behaviour (signed overflow, misaligned or null
  access, out-of-range shift) aborts the run with the sanitizer's report.
* The SciDAC writer and reader of the library (chroma_amd/csrc/qio.cpp), compiled for the host with
  the same sanitizers (tests/native/Makefile), run round trips and their error paths (wrong dims, a
  cut file, a flipped byte, a missing file).
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_TESTS = ["tests/test_oracle.py", "tests/test_oracle_force.py", "tests/test_oracle_krylov.py",
                "tests/test_oracle_learned.py"]


def _make(path, target=None):
    cmd = ["make", "-s", "-C", path] + ([target] if target else [])
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout + p.stderr


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, timeout=60)
    path = p.stdout.strip()
    if p.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("gcc's libasan is not installed")
    return path


def _report(out):
    return "ERROR: AddressSanitizer" in out or "runtime error:" in out or "ERROR: LeakSanitizer" in out


def test_oracle_under_asan_ubsan():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    _make(os.path.join(ROOT, "oracle"), "san")
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle_san.so")
    env = dict(os.environ, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", CMX_ORACLE_LIB=lib)
    # the sanitizer build is the library the tests load
    p = subprocess.run([sys.executable, "-c", "import oracle; print(oracle.lib()._name)"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0 and p.stdout.strip() == lib, p.stdout + p.stderr
    # --noconftest: the GPU probe in conftest imports torch, which is not what is checked here
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "--noconftest", "-p", "no:cacheprovider"] +
                       ORACLE_TESTS, cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = p.stdout + p.stderr
    assert p.returncode == 0 and not _report(out), out[-4000:]
    assert " passed" in out


def test_scidac_under_asan_ubsan(tmp_path):
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("ROCm clang is not installed")
    _make(os.path.join(ROOT, "tests", "native"))
    exe = os.path.join(ROOT, "tests", "native", "_san", "scidac_san")
    p = subprocess.run([exe, str(tmp_path)], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"),
                       capture_output=True, text=True, timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0 and not _report(out), out[-4000:]
    assert "PASSED: 0 failure(s)" in out
    assert out.count("round trip") == 8
