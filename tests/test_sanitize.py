"""The oracle and the library's host code built with AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/sanitize_check.cpp, `make -C oracle sanitize`), run over the
golden fixtures: no invalid access, leak or undefined behaviour."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "build", "sanitize_check")


@pytest.fixture(scope="module")
def exe():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    return EXE


def test_sanitizers_are_live(exe):
    r = subprocess.run([exe, "--planted-overflow"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr


def test_host_code_and_oracle_under_sanitizers(exe, tmp_path):
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden"), str(tmp_path / "scratch.coef")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "lyapunov: done" in r.stdout
