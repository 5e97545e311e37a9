"""The host-side parsers, tables and the oracle under the address and undefined-behaviour sanitizers (CPU).

* `make -C dpgsolver_amd/csrc host_check` builds dpgsolver_amd/csrc/host_check.cpp with the library's host
  sources (gmsh reader, restart writer/reader, face permutations, TRI tables) under g++ -fsanitize=address,undefined
  and runs it on the golden meshes: every mesh at p = 1 and 3, the restart round trip, truncated copies of each
  mesh (which must come back as a status code), and arguments each entry point must refuse.
* `make -C oracle sanitize_check` builds oracle/sanitize_check.c with the oracle under the same sanitizers and runs
  every residual entry point on small curved periodic QUAD and HEX meshes, plus the point functions over a spread of
  states (every Riemann boundary branch).
Any out-of-bounds access, leak, use after free or undefined behaviour aborts the program, and the test fails.
"""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not available")


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300, **kw)


def test_host_parsers_under_sanitizers(tmp_path):
    _need("g++")
    b = _run(["make", "-C", os.path.join(ROOT, "dpgsolver_amd", "csrc"), "host_check"])
    assert b.returncode == 0, b.stdout + b.stderr
    meshes = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mesh", "*.msh")))
    assert meshes
    r = _run([os.path.join(ROOT, "dpgsolver_amd", "csrc", "_build", "host_check"), str(tmp_path)] + meshes,
             env={**os.environ, "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host_check: ok" in r.stdout


def test_oracle_under_sanitizers():
    _need("gcc")
    b = _run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize_check"])
    assert b.returncode == 0, b.stdout + b.stderr
    r = _run([os.path.join(ROOT, "oracle", "_san", "sanitize_check")],
             env={**os.environ, "UBSAN_OPTIONS": "print_stacktrace=1", "OMP_NUM_THREADS": "3"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize_check: ok" in r.stdout
