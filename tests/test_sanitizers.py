"""Sanitized host builds: the oracle (oracle/ugacc_oracle.c, driven by
tests/sanitize_oracle.c through CCSD, the three (T) forms, HBAR, Lambda,
densities, both (T) gradient forms and the linear response of a symmetric
and an antisymmetric perturbation) and the front end's integral code
(ugacc_amd/chem/integrals.cpp, driven by tests/sanitize_chem.cpp through
every integral class with s, p and d shells), compiled with
-fsanitize=address,undefined (gcc / g++, host code only; OpenMP pragmas off)
and run with leak detection.  Any out-of-bounds access, use after free,
leak or undefined behaviour fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-Wall", "-Wno-unknown-pragmas"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build_and_run(tmp_path, compiler, std, sources, libs):
    exe = str(tmp_path / "san")
    cmd = [compiler, std] + FLAGS + [os.path.join(ROOT, s) for s in sources] + libs + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr and ("cannot find" in r.stderr or "unrecognized" in r.stderr):
        pytest.skip("the host compiler has no sanitizer runtime: " + r.stderr.strip()[:200])
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert r.stdout.startswith("OK"), r.stdout
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_oracle_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "gcc", "-std=c11", ["oracle/ugacc_oracle.c", "tests/sanitize_oracle.c"], ["-lm"])
    assert "ccsd rc 0" in out and "lambda rc 0" in out and "pert rc 0" in out, out


def test_integrals_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    _build_and_run(tmp_path, "g++", "-std=c++17", ["ugacc_amd/chem/integrals.cpp", "tests/sanitize_chem.cpp"], [])
