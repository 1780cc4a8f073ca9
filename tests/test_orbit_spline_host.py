"""Host-side Spline operations of Orbit objects (agama_amd.orbitobj.QuinticPieces: evaluation up
to the third derivative, integrals, extrema, roots; the reference's Spline_value /
Spline_integrate / Spline_roots / Spline_extrema, src/interface_python.cpp:6159-6330, on
QuinticSpline, src/math_spline.cpp:1391-1463) pinned by a global quintic polynomial, which the
quintic spline through its values and first two derivatives reproduces exactly."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from agama_amd.orbitobj import QuinticPieces  # noqa: E402

P = np.polynomial.Polynomial([0.3, -1.2, 0.5, 0.8, -0.3, 0.02])
T = np.sort(np.concatenate([[0.0, 5.0], np.random.default_rng(1).uniform(0.1, 4.9, 9)]))


def spline(poly=P, t=T):
    return QuinticPieces(t, poly(t), poly.deriv(1)(t), poly.deriv(2)(t))


def real_roots(poly, lo, hi):
    r = poly.roots()
    r = np.sort(r[np.abs(r.imag) < 1e-10].real)
    return r[(r >= lo) & (r <= hi)]


def test_values_and_derivatives():
    q = spline()
    x = np.linspace(0, 5, 2001)
    scale = np.max(np.abs(P(x)))
    for der, tol in ((0, 1e-13), (1, 1e-11), (2, 1e-8), (3, 1e-6)):
        assert np.max(np.abs(q(x, der) - P.deriv(der)(x))) <= tol * max(scale, 1), der
    # linear continuation beyond the end nodes, zero second and third derivatives there
    out = np.array([-1.5, 6.25])
    assert np.allclose(q(out), [P(0) - 1.5 * P.deriv()(0), P(5) + 1.25 * P.deriv()(5)], rtol=1e-13)
    assert np.allclose(q(out, 1), [P.deriv()(0), P.deriv()(5)], rtol=1e-13)
    assert np.all(q(out, 2) == 0) and np.all(q(out, 3) == 0)


@pytest.mark.parametrize("n", [0, 1, 2, 5])
def test_integrals_with_monomial(n):
    q = spline()
    F = (P * np.polynomial.Polynomial([0] * n + [1])).integ()
    for a, b in ((0.7, 4.2), (0.0, 5.0), (1.3, 1.31)):
        exact = F(b) - F(a)
        assert q.integrate(a, b, n) == pytest.approx(exact, rel=1e-12, abs=1e-12)
        assert q.integrate(b, a, n) == pytest.approx(-exact, rel=1e-12, abs=1e-12)
    # zero outside the grid
    assert q.integrate(-3.0, 8.0, n) == pytest.approx(F(5.0) - F(0.0), rel=1e-12)
    assert q.integrate(6.0, 9.0, n) == 0


def test_integral_of_a_product():
    Q = np.polynomial.Polynomial([1.0, 0.2, -0.1, 0.03])
    t2 = np.linspace(-1, 3.5, 7)        # another grid, overlapping [0, 3.5]
    prod = (P * Q).integ()
    assert spline().integrate(0.5, 4.5, spline(Q, t2)) == pytest.approx(prod(3.5) - prod(0.5), rel=1e-12)


def test_extrema_and_roots():
    q = spline()
    ex = q.extrema()
    interior = real_roots(P.deriv(), 0, 5)
    assert ex[0] == 0 and ex[-1] == 5
    assert np.allclose(ex[1:-1], interior, rtol=1e-10, atol=1e-10)
    for y in (0.0, 0.5, -1.0, 10.0):
        assert np.allclose(q.roots(y), real_roots(P - y, 0, 5), rtol=1e-12, atol=1e-12), y
    # a sub-interval
    assert np.allclose(q.roots(0.0, 0.5, 4.0), real_roots(P, 0.5, 4.0), rtol=1e-12)
    sub = q.extrema(1.0, 3.0)
    assert sub[0] == 1.0 and sub[-1] == 3.0


def test_stationary_inflection_is_not_an_extremum():
    C = np.polynomial.Polynomial([-8.0, 12.0, -6.0, 1.0])     # (x - 2)^3: monotonic
    q = spline(C, np.linspace(0, 5, 6))
    assert np.allclose(q.extrema(), [0.0, 5.0])
    assert np.allclose(q.roots(0.0), [2.0], atol=1e-5)         # a triple root: flat


@pytest.mark.parametrize("sigma", [0.05, 0.4, 2.0])
def test_gaussian_convolution(sigma):
    from scipy.integrate import quad
    q = spline()
    xs = np.array([-1.0, 0.3, 2.5, 4.9, 5.6])
    got = q.convolve(xs, sigma)
    for x, g in zip(xs, got):
        f = lambda y: P(y) * np.exp(-0.5 * ((x - y) / sigma) ** 2) / (np.sqrt(2 * np.pi) * sigma)  # noqa: E731
        ref = quad(f, 0.0, 5.0, points=[x] if 0 < x < 5 else None, epsabs=1e-13, epsrel=1e-12, limit=200)[0]
        assert g == pytest.approx(ref, rel=1e-10, abs=1e-13)


def test_convolution_with_another_spline():
    from scipy.integrate import quad
    Q = np.polynomial.Polynomial([1.0, -0.5, 0.25])
    tq = np.linspace(-1.0, 1.5, 6)
    q, k = spline(), spline(Q, tq)
    for x in (-0.5, 0.7, 3.0, 6.0, 7.0):
        lo, hi = max(0.0, x - 1.5), min(5.0, x + 1.0)
        ref = quad(lambda y: P(y) * Q(x - y), lo, hi, epsabs=1e-14, epsrel=1e-13)[0] if lo < hi else 0.0
        assert q.convolve([x], k)[0] == pytest.approx(ref, rel=1e-11, abs=1e-13)
