"""The C++ layer include/agama_amd.hpp (the reference's BasePotential::eval / OrbitIntParams /
batched OrbitIntegrator<Car> surface over the C ABI), exercised by a compiled program
(tests/cpp/adapter_check.cpp).  CPU: it compiles against the header and the library and its
error paths throw std::runtime_error with the library's message.  GPU: single-point eval()
equals evalMany() of the same kernel bit for bit, NULL outputs are skipped, integrateOrbits()
equals the last sample of integrateTrajectories(), and the values equal the oracle's (potential,
gradient and Hessian within 1e-12 of the largest component; orbit final positions within 1e-7
relative for the median, 3 circular periods at accuracy 1e-8)."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
BINARY = os.path.join(ROOT, "tests", "cpp", "build", "adapter_check")


def test_adapter_builds_and_reports_errors():
    from tests.cpp.build_adapter_check import build
    exe = build()
    out = subprocess.run([exe, "errors"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "errors ok" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_adapter_values_match_oracle(tmp_path):
    from oracle.pyoracle import OraclePotential
    assert os.path.exists(BINARY), "build it first: python tests/cpp/build_adapter_check.py"
    path = str(tmp_path / "values.bin")
    print("case: adapter_check values (2000 points, 64 orbits x 3 periods)", flush=True)
    out = subprocess.run([BINARY, "values", COEF, path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "values ok" in out.stdout, out.stdout + out.stderr
    raw = np.fromfile(path, dtype=np.float64)
    n, m = 2000, 64
    sizes = [n * 3, n, n * 3, n * 6, m * 6, m, m * 6, m]
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    pts, phi, g, h = parts[0].reshape(n, 3), parts[1], parts[2].reshape(n, 3), parts[3].reshape(n, 6)
    ic, T, fin = parts[4].reshape(m, 6), parts[5], parts[6].reshape(m, 6)
    nsteps = parts[7].view(np.int64)
    orc = OraclePotential(COEF)
    ophi, og, oh = orc.eval_hess(pts)
    assert np.max(np.abs(phi / ophi - 1)) <= 1e-12
    assert np.max(np.abs(g - og).max(axis=1) / np.abs(og).max(axis=1)) <= 1e-12
    assert np.max(np.abs(h - oh).max(axis=1) / np.abs(oh).max(axis=1)) <= 1e-12
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8)
    err = np.linalg.norm(fin[:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    print(f"adapter orbits: position err median {np.median(err):.1e}, identical step counts "
          f"{np.mean(nsteps == ref['nsteps']):.3f}")
    assert np.median(err) <= 1e-7 and np.mean(nsteps == ref["nsteps"]) >= 0.9
