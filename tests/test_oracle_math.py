"""Pins of the oracle's scalar helpers, Legendre functions, trig recurrence and 2-D quintic
spline (oracle/agama_oracle.cpp) against libm, scipy closed forms and exact reproduction."""
import ctypes
import math

import numpy as np
import pytest
from scipy import special

from oracle.pyoracle import lib as orc


def sincos(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    orc.orc_sincos(x, ctypes.addressof(s), ctypes.addressof(c))
    return s.value, c.value


def test_sincos_accuracy_and_exact_quadrants():
    xs = np.concatenate([np.linspace(-50, 50, 20001), [1e-300, -1e-8, 1e-8, 3.0, 1e3]])
    for x in xs:
        s, c = sincos(float(x))
        assert abs(s - math.sin(x)) < 5e-16 * max(1, abs(x)) and abs(c - math.cos(x)) < 5e-16 * max(1, abs(x))
    for k in range(-8, 9):   # deliberate exact zeros at multiples of pi/2 (math_core.cpp:176-178)
        s, c = sincos(k * math.pi / 2)
        assert (s == 0) if k % 2 == 0 else (c == 0)


def test_atan2_accuracy_and_signed_zeros():
    rng = np.random.default_rng(0)
    for y, x in rng.normal(size=(20000, 2)) * 10 ** rng.uniform(-5, 5, size=(20000, 1)):
        assert abs(orc.orc_atan2(y, x) - math.atan2(y, x)) <= 4.5e-16 * max(1.0, abs(math.atan2(y, x)))
    assert orc.orc_atan2(0.0, -1.0) == math.pi and orc.orc_atan2(-0.0, -1.0) == -math.pi
    assert orc.orc_atan2(0.0, 1.0) == 0.0 and math.copysign(1, orc.orc_atan2(-0.0, 1.0)) == -1


def test_binsearch_edges():
    arr = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    p = arr.ctypes.data
    assert orc.orc_binsearch(-0.1, p, 5) == -1
    assert orc.orc_binsearch(1.5, p, 5) == 4
    assert orc.orc_binsearch(1.0, p, 5) == 3        # exactly at the end: last interval
    assert orc.orc_binsearch(0.0, p, 5) == 0
    assert [orc.orc_binsearch(x, p, 5) for x in (0.05, 0.1, 0.29, 0.3, 0.7)] == [0, 1, 1, 2, 3]
    assert orc.orc_binsearch(float("nan"), p, 5) == -1


@pytest.mark.parametrize("wm1", [0, 1])
def test_lambert_w_both_branches(wm1):
    for x in np.concatenate([-np.logspace(-12, math.log10(1 / math.e) - 1e-12, 300), [-1 / math.e + 1e-15]]):
        w = orc.orc_lambertw(float(x), wm1)
        # defining identity W exp(W) = x, to rounding (exp amplifies the rounding of W by |W|)
        assert abs(w * math.exp(w) - x) <= (3 + abs(w)) * 2.3e-16 * abs(x)
        assert (w <= -1) if wm1 else (w >= -1)
        # away from the branch point (where scipy's value itself loses digits) agree with scipy
        if math.e * x + 1 > 1e-6:
            expected = special.lambertw(x, k=-1 if wm1 else 0).real
            assert abs(w - expected) <= 4e-15 * abs(expected)


@pytest.mark.parametrize("n", [1, 2, 3, 9, 10, 13, 15, 17, 23, 33])
def test_gauss_legendre_nodes(n):
    nodes, weights = np.empty(n), np.empty(n)
    orc.orc_gl_nodes(n, nodes.ctypes.data, weights.ctypes.data)
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(nodes - x)) < 2e-15 and np.max(np.abs(weights - w)) < 4e-15
    for k in range(2 * n):   # exact for polynomials of degree < 2n
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(weights * nodes ** k) - exact) < 1e-13


@pytest.mark.parametrize("m", [0, 1, 2, 3, 6, 8, 16])
def test_normalised_legendre_vs_closed_form(m):
    lmax = 18
    for tau in np.linspace(-0.999, 0.999, 41):
        P, dP = np.empty(lmax - m + 1), np.empty(lmax - m + 1)
        orc.orc_sphharm_array(lmax, m, float(tau), P.ctypes.data, dP.ctypes.data)
        ct, st = 2 * tau / (1 + tau * tau), (1 - tau * tau) / (1 + tau * tau)
        theta = math.atan2(st, ct)
        for l in range(m, lmax + 1):
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))
            val = norm * special.lpmv(m, l, ct)
            assert abs(P[l - m] - val) < 1e-11 * max(1, abs(val)), (l, m, tau)
            h = 1e-6
            fd = norm * (special.lpmv(m, l, math.cos(theta + h)) - special.lpmv(m, l, math.cos(theta - h))) / (2 * h)
            assert abs(dP[l - m] - fd) < 1e-6 * max(1, abs(fd)), (l, m, tau)


def test_trig_multi_angle():
    for phi in np.linspace(-7, 7, 101):
        out = np.empty(2 * 16)
        orc.orc_trig_multi_angle(float(phi), 16, 1, out.ctypes.data)
        k = np.arange(1, 17)
        tol = 1e-15 * k * (1 + abs(phi))   # recurrence error grows ~k; k*phi itself is rounded
        assert np.all(np.abs(out[:16] - np.cos(k * phi)) < tol)
        assert np.all(np.abs(out[16:] - np.sin(k * phi)) < tol)


def _spline(xg, yg, f, fx, fy, fxy, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(pts)
    z, zx, zy = np.empty(n), np.empty(n), np.empty(n)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (xg, yg, f, fx, fy, fxy)]
    rc = orc.orc_spline2d_eval(len(xg), len(yg), *[a.ctypes.data for a in args], n, pts.ctypes.data,
                               z.ctypes.data, zx.ctypes.data, zy.ctypes.data)
    assert rc == 0
    return z, zx, zy


def test_spline2d_node_exactness_and_nan_outside():
    rng = np.random.default_rng(1)
    xg = np.cumsum(rng.uniform(0.2, 1.0, 9))
    yg = np.cumsum(rng.uniform(0.1, 0.5, 7))
    f, fx, fy, fxy = (rng.normal(size=(9, 7)) for _ in range(4))
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    z, zx, zy = _spline(xg, yg, f, fx, fy, fxy, np.stack([X.ravel(), Y.ravel()], 1))
    assert np.array_equal(z, f.ravel()) and np.array_equal(zx, fx.ravel()) and np.array_equal(zy, fy.ravel())
    z, _, _ = _spline(xg, yg, f, fx, fy, fxy, [[xg[0] - 0.1, yg[2]], [xg[3], yg[-1] + 0.1]])
    assert np.all(np.isnan(z))


def test_spline2d_reproduces_bicubic():
    rng = np.random.default_rng(2)
    c = rng.normal(size=(4, 4))
    xg = np.cumsum(rng.uniform(0.2, 1.0, 8))
    yg = np.cumsum(rng.uniform(0.2, 1.0, 6))
    px = np.polynomial.polynomial
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    f = px.polyval2d(X, Y, c)
    fx = px.polyval2d(X, Y, px.polyder(c, axis=0))
    fy = px.polyval2d(X, Y, px.polyder(c, axis=1))
    fxy = px.polyval2d(X, Y, px.polyder(px.polyder(c, axis=0), axis=1))
    pts = np.stack([rng.uniform(xg[0], xg[-1], 500), rng.uniform(yg[0], yg[-1], 500)], 1)
    z, zx, zy = _spline(xg, yg, f, fx, fy, fxy, pts)
    scale = np.max(np.abs(f))
    assert np.max(np.abs(z - px.polyval2d(pts[:, 0], pts[:, 1], c))) < 1e-11 * scale
    assert np.max(np.abs(zx - px.polyval2d(pts[:, 0], pts[:, 1], px.polyder(c, axis=0)))) < 1e-10 * scale
    assert np.max(np.abs(zy - px.polyval2d(pts[:, 0], pts[:, 1], px.polyder(c, axis=1)))) < 1e-10 * scale
