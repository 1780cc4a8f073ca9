"""Pins of the oracle's second derivatives (oracle/agama_oracle.cpp), restating
BasePotential::eval with deriv2 for the Multipole path: QuinticSpline2d::evalDeriv's z_xx, z_xy,
z_yy (src/math_spline.cpp:2535-2597), sphHarmArray's d2P (src/math_sphharm.cpp:63-132),
MultipoleInterp2d::evalCyl / PowerLawMultipole::evalCyl with hess
(src/potential_multipole.cpp:1969-2080, 666-848, 1600-1658), the coordinate Hessian transforms
(src/coord.cpp:94-125, 700-720, 779-799, 923-956) and the NFW deriv2 branches
(src/potential_analytic.cpp:55-75).
Checks: exact reproduction of a bicubic by the spline's second derivatives; the Legendre ODE
and finite differences of the (pinned) first derivatives; continuity of the pole branches;
fourth-order finite differences of the (pinned) gradient in every region of the expansion;
the NFW closed form and its Laplacian 4 pi rho."""
import math
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN
from oracle.pyoracle import lib as orc, OraclePotential

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
BAR = os.path.join(GOLDEN, "ferrers_bar_lmax16.coef")


def test_spline2d_second_derivatives_reproduce_bicubic():
    rng = np.random.default_rng(3)
    c = rng.normal(size=(4, 4))
    xg = np.cumsum(rng.uniform(0.2, 1.0, 8))
    yg = np.cumsum(rng.uniform(0.2, 1.0, 6))
    px = np.polynomial.polynomial
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    der = lambda cc, ax: px.polyder(cc, axis=ax)
    arrs = [px.polyval2d(X, Y, a) for a in (c, der(c, 0), der(c, 1), der(der(c, 0), 1))]
    pts = np.stack([rng.uniform(xg[0], xg[-1], 500), rng.uniform(yg[0], yg[-1], 500)], 1)
    pts[:5] = [[xg[2], yg[3]], [xg[-1], yg[-1]], [xg[0], yg[1]], [xg[4], 0.5 * (yg[1] + yg[2])],
               [0.5 * (xg[0] + xg[1]), yg[-1]]]
    n = len(pts)
    out = [np.empty(n) for _ in range(6)]
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in [xg, yg] + arrs]
    rc = orc.orc_spline2d_eval2(len(xg), len(yg), *[a.ctypes.data for a in args], n, pts.ctypes.data,
                                *[o.ctypes.data for o in out])
    assert rc == 0
    z, zx, zy, zxx, zxy, zyy = out
    scale = np.max(np.abs(arrs[0]))
    x, y = pts[:, 0], pts[:, 1]
    assert np.max(np.abs(zxx - px.polyval2d(x, y, der(der(c, 0), 0)))) < 1e-9 * scale
    assert np.max(np.abs(zxy - px.polyval2d(x, y, der(der(c, 0), 1)))) < 1e-9 * scale
    assert np.max(np.abs(zyy - px.polyval2d(x, y, der(der(c, 1), 1)))) < 1e-9 * scale
    assert np.max(np.abs(zx - px.polyval2d(x, y, der(c, 0)))) < 1e-10 * scale


def _sph2(lmax, m, tau):
    k = lmax - m + 1
    P, dP, d2P = np.empty(k), np.empty(k), np.empty(k)
    orc.orc_sphharm_array2(lmax, m, float(tau), P.ctypes.data, dP.ctypes.data, d2P.ctypes.data)
    return P, dP, d2P


def _tau(theta):
    # inverse of cos(theta) = 2 tau / (1 + tau^2), sin(theta) = (1 - tau^2) / (1 + tau^2)
    return math.tan(math.pi / 4 - theta / 2)


@pytest.mark.parametrize("m", [0, 1, 2, 3, 5, 8])
def test_legendre_second_derivative(m):
    lmax = 16
    for theta in np.linspace(0.02, math.pi - 0.02, 37):
        P, dP, d2P = _sph2(lmax, m, _tau(theta))
        ct, st = math.cos(theta), math.sin(theta)
        l = np.arange(m, lmax + 1)
        # Legendre ODE in theta: P'' = -cot(theta) P' - (l(l+1) - m^2/sin^2) P
        ode = -ct / st * dP - (l * (l + 1) - m * m / st ** 2) * P
        scale = np.maximum(1, np.abs(d2P))
        assert np.max(np.abs(d2P - ode) / scale) < 1e-9, theta
        # central difference of the first derivatives
        h = 1e-5
        fd = (_sph2(lmax, m, _tau(theta + h))[1] - _sph2(lmax, m, _tau(theta - h))[1]) / (2 * h)
        assert np.max(np.abs(d2P - fd) / scale) < 1e-6, theta


@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_legendre_second_derivative_pole_branches(m):
    """below sin(theta) = 1e-8 the reference switches to pole forms (and for m=1 to its own
    recurrence): they must continue the general expression.  Near a pole d2P is even in the
    distance to the pole for even m and odd (a slope) for odd m."""
    lmax = 16
    near = 1e-3
    for north in (True, False):
        th_near = near if north else math.pi - near
        _, _, ref = _sph2(lmax, m, _tau(th_near))
        for eps in (0.0, 5e-9):
            th = eps if north else math.pi - eps
            _, _, d2P = _sph2(lmax, m, 1.0 if (north and eps == 0) else -1.0 if eps == 0 else _tau(th))
            expect = ref if m % 2 == 0 else ref * (eps / near)
            scale = np.maximum(1, np.abs(ref))
            assert np.max(np.abs(d2P - expect) / scale) < 1e-2, (m, north, eps)
            if m % 2 == 1 and eps > 0:   # the slope itself, relative
                big = np.abs(ref) > 1e-3
                assert np.max(np.abs(d2P[big] / expect[big] - 1)) < 1e-2, (m, north)


def _fd_hessian(orc_pot, pts, rel=2e-5):
    """fourth-order central differences of the oracle's gradient"""
    pts = np.asarray(pts, dtype=np.float64)
    n = len(pts)
    h = rel * np.linalg.norm(pts, axis=1)
    H = np.empty((n, 3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        g = [orc_pot.eval(pts + s * h[:, None] * e)[1] for s in (2, 1, -1, -2)]
        H[:, :, k] = (-g[0] + 8 * g[1] - 8 * g[2] + g[3]) / (12 * h[:, None])
    return H


def _as_matrix(h6):
    d = np.empty((len(h6), 3, 3))
    d[:, 0, 0], d[:, 1, 1], d[:, 2, 2] = h6[:, 0], h6[:, 1], h6[:, 2]
    d[:, 0, 1] = d[:, 1, 0] = h6[:, 3]
    d[:, 1, 2] = d[:, 2, 1] = h6[:, 4]
    d[:, 0, 2] = d[:, 2, 0] = h6[:, 5]
    return d


def _points(rng, n, rmin, rmax):
    r = np.exp(rng.uniform(np.log(rmin), np.log(rmax), n))
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]


@pytest.mark.parametrize("coef,nfw", [(COEF, None), (BAR, (1.0, 20.0))])
def test_hessian_matches_finite_differences_of_gradient(coef, nfw):
    """interior spline (both the m=0 log-scaling and the ratio terms), inner and outer
    power-law extrapolation, and the composite with an NFW halo"""
    o = OraclePotential(coef, nfw=nfw)
    rng = np.random.default_rng(17)
    pts = np.concatenate([_points(rng, 300, 2e-4, 5e3),
                          [[0.3, 0.2, 0.5], [1.0, 1e-3, 0.0], [2.0, -0.5, 1e-4], [1e-3, 2e-3, 0.5]]])
    phi, g, h6 = o.eval_hess(pts)
    phi2, g2 = o.eval(pts)
    assert np.array_equal(phi, phi2) and np.array_equal(g, g2)   # same path as the pinned gradient
    H = _as_matrix(h6)
    F = _fd_hessian(o, pts)
    err = np.max(np.abs(H - F), axis=(1, 2)) / np.max(np.abs(F), axis=(1, 2))
    print(f"Hessian vs finite differences: median {np.median(err):.1e}, max {err.max():.1e}")
    assert np.median(err) < 1e-8
    assert err.max() < 1e-5


def test_nfw_hessian_closed_form_and_laplacian():
    """the halo's part of the composite: Phi'' closed form (general and small-r rational
    branches) and trace = 4 pi rho_NFW"""
    M, a = 1.0, 20.0
    bar = OraclePotential(BAR)
    both = OraclePotential(BAR, nfw=(M, a))
    rng = np.random.default_rng(5)
    r = np.concatenate([np.geomspace(1e-3, 2e3, 300), [a * 0.0099, a * 0.0101]])
    d = rng.normal(size=(len(r), 3))
    pts = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    hn = both.eval_hess(pts)[2] - bar.eval_hess(pts)[2]
    q = r / a
    d1 = M * (np.log1p(q) / r ** 2 - 1 / (r * (r + a)))
    d2 = -M * (2 * np.log1p(q) / r ** 3 - (2 * a + 3 * r) / (r ** 2 * (a + r) ** 2))
    u = pts / r[:, None]
    ref = d2[:, None, None] * u[:, :, None] * u[:, None, :] + \
        (d1 / r)[:, None, None] * (np.eye(3)[None] - u[:, :, None] * u[:, None, :])
    H = _as_matrix(hn)
    err = np.max(np.abs(H - ref), axis=(1, 2)) / np.max(np.abs(ref), axis=(1, 2))
    # the rational form below r/a = 0.01 carries the 10 digits the reference quotes; the sum
    # subtracts the bar's part
    assert np.max(err[q > 0.02]) < 1e-9 and np.max(err) < 1e-7
    rho = M / (4 * np.pi * a ** 3) / (q * (1 + q) ** 2)
    lap = hn[:, 0] + hn[:, 1] + hn[:, 2]
    assert np.max(np.abs(lap / (4 * np.pi * rho) - 1)[q > 0.02]) < 1e-8
