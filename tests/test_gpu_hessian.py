"""GPU parity of the Hessian evaluation (evalHessKernel: BasePotential::eval with deriv2 for the
Multipole path) against the CPU oracle's second derivatives, on the triaxial Dehnen Multipole
(lmax=8) and on the Ferrers bar (lmax=16) + NFW composite, at points drawn across the inner
power-law, spline and outer power-law regions; and the Python surfaces eval(der=True) and
forceDeriv (force derivatives = -Hessian, order xx, yy, zz, xy, yz, zx).
Bar: potential and gradient 1e-12 relative (as for the first-derivative path); each Hessian
within 1e-11 of its largest component."""
import ctypes
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
BAR = os.path.join(GOLDEN, "ferrers_bar_lmax16.coef")


def _points(seed, n):
    g = np.random.Generator(np.random.Philox(seed))
    r = np.exp(g.uniform(np.log(1e-4), np.log(5e3), n))
    d = g.normal(size=(n, 3))
    pts = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    pts[:8] = [[0, 0, 0.5], [0, 0, -2], [1, 0, 0], [0, 1, 0], [1e-5, 0, 0], [800, 0, 0], [0.3, -0.2, 0.1],
               [2.9, 0.1, 0.0]]
    return pts


@pytest.mark.parametrize("case", ["dehnen", "bar_nfw"])
def test_hessian_matches_oracle(case):
    import agama_amd
    from oracle.pyoracle import OraclePotential
    if case == "dehnen":
        gpu, orc = agama_amd.Potential(file=COEF), OraclePotential(COEF)
    else:
        gpu = agama_amd.Potential(dict(type="Multipole", file=BAR), dict(type="NFW", mass=1, scaleRadius=20))
        orc = OraclePotential(BAR, nfw=(1.0, 20.0))
    pts = _points(91 if case == "dehnen" else 92, 200_000)
    print(f"case: hessian {case} 200k points", flush=True)
    phi, grad, hess = gpu._eval_hess(pts)
    ophi, og, oh = orc.eval_hess(pts)
    ep = np.abs(phi / ophi - 1)
    eg = np.linalg.norm(grad - og, axis=1) / np.linalg.norm(og, axis=1)
    eh = np.max(np.abs(hess - oh), axis=1) / np.max(np.abs(oh), axis=1)
    print(f"{case}: potential {ep.max():.2e}, gradient {eg.max():.2e}, hessian median {np.median(eh):.2e} "
          f"max {eh.max():.2e}")
    assert ep.max() <= 1e-12 and eg.max() <= 1e-12
    assert eh.max() <= 1e-11
    # the gradient of the Hessian path (node records) is that of the first-derivative path (the
    # per-cell polynomial form of the same splines), to rounding
    phi1, grad1 = gpu._eval(pts, True, True)
    assert np.max(np.abs(phi1 / phi - 1)) <= 1e-12
    assert np.max(np.linalg.norm(grad1 - grad, axis=1) / np.linalg.norm(grad, axis=1)) <= 1e-12


def test_python_and_c_surfaces():
    import agama_amd
    from agama_amd import lib
    gpu = agama_amd.Potential(file=COEF)
    pts = _points(93, 1000)
    phi, grad, hess = gpu._eval_hess(pts)
    p2, acc, der = gpu.eval(pts, pot=True, acc=True, der=True)
    assert np.array_equal(p2, phi) and np.array_equal(acc, -grad) and np.array_equal(der, -hess)
    f, d = gpu.forceDeriv(pts[:5])
    assert f.shape == (5, 3) and d.shape == (5, 6) and np.array_equal(d, -hess[:5])
    d1 = gpu.eval(pts[7], der=True)
    assert d1.shape == (6,) and np.array_equal(d1, -hess[7])
    # agama_evalPotential with deriv2 (interface_c.cpp:68-98): +Hessian of Phi
    x = np.ascontiguousarray(pts[6])
    g3, h6 = np.empty(3), np.empty(6)
    v = lib.agama_evalPotential(gpu.handle, x.ctypes.data, 0.0, g3.ctypes.data, h6.ctypes.data)
    assert v == phi[6] and np.array_equal(g3, grad[6]) and np.array_equal(h6, hess[6])
    # FP32 tables do not provide second derivatives
    rc = lib.agx_eval_hess_many(gpu.handle, 1, x.ctypes.data, 0.0, None, None, h6.ctypes.data, 1)
    assert rc == 3
