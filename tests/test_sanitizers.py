"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the CPU code (no GPU).

tests/sanitize/Makefile builds, with the host compilers:
  oracle_san  every entry point of the FP64 CPU oracle (tree, lists, ACA,
              stored / skeleton / sampled operators, matvec, GMRES, CPQR,
              quadrature) on 1-D to 4-D random, grid and clustered points;
  host_san    the library's host-only code: tree and list construction
              (tree.cpp) checked for structural invariants, and the
              reference's vector files (vecio.cpp) including truncated,
              lying and foreign files.
A sanitizer report aborts the program (-fno-sanitize-recover=all), so a
clean exit with the final "run clean" line is the pass condition.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


def _have_sanitizer_runtime():
    if not shutil.which("gcc") or not shutil.which("g++"):
        return False
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    return r.returncode == 0 and os.path.sep in r.stdout.strip()


pytestmark = pytest.mark.skipif(not _have_sanitizer_runtime(), reason="gcc with libasan not available")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return os.path.join(SAN, "_build")


def _run(cmd):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    return subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)


def test_oracle_under_asan_ubsan(built):
    r = _run([os.path.join(built, "oracle_san")])
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "oracle sanitizer run clean" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out


def test_host_tree_and_vector_files_under_asan_ubsan(built, tmp_path):
    r = _run([os.path.join(built, "host_san"), str(tmp_path)])
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "host sanitizer run clean" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out
