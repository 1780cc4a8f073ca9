"""Pins of the oracle's Multipole evaluation against mathematics (CPU only).

The oracle restates MultipoleInterp2d::evalCyl, PowerLawMultipole::evalCyl and the Fourier
synthesis (src/potential_multipole.cpp:1538-1548, 1585-1658, 1969-2080); the GPU kernels are
checked against it elsewhere.  Here nothing of the oracle is retyped: every expected value comes
from the coefficient file or from an independent computation.

  * node lattice: the 2-D spline passes through its nodes, whose values are the spherical-
    harmonic synthesis of the file's #Phi block (and whose ln r derivatives that of the #dPhi/dr
    block) at (r_k, tau_j): Phi = sum_lm C_lm(r_k) sqrt((2l+1)(l-|m|)!/(l+|m|)!) P_l^|m|(cos theta)
    {1, sqrt2 cos m phi, sqrt2 sin |m| phi} (src/math_sphharm.cpp:22-61, 386-414), with scipy's
    lpmv (Condon-Shortley phase, as legendrePmm);
  * the gradient equals fourth-order central differences of the potential, inside the grid and
    in both power-law regions;
  * potential and gradient are continuous across the branch radii r_min(1+eps) and
    r_max(1-eps), eps = cbrt(DBL_EPSILON) (src/potential_multipole.cpp:59, 1538-1548) - except
    where the reference itself drops terms: the outer extrapolation keeps only the harmonics
    whose U coefficient is non-zero (its index set comes from getIndicesFromCoefs(U),
    src/potential_multipole.cpp:1590), and for the compact Ferrers bar every m != 0 term takes
    the round-off branch of computeExtrapolationCoefs (U = 0, :579-586); the jump there equals
    the m != 0 synthesis at r_max;
  * far field: Phi r -> -M and |grad Phi| r^2 -> M for the finite-mass models;
  * the product's Poisson route (computePotentialCoefsFromSource, src/potential_multipole.cpp:
    1013-1114, in agx_build_coef_file) equals an independent quadrature of the Green's-function
    integrals Phi_lm(r) = -4 pi/(2l+1) [r^-(l+1) int_0^r rho_lm s^(l+2) ds + r^l int_r^inf rho_lm
    s^(1-l) ds], with rho_lm from a 96 x 96 Gauss-Legendre x trapezoid angular quadrature (beyond
    the aliasing of the product's order-14 transform for this model) and composite Gauss-Legendre
    in ln s; this pins the forward spherical-harmonic transform and the radial recurrences of
    every (l, m) term, not only l = 0.
"""
import math
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

FIXTURES = ["dehnen_triax_lmax8.coef", "disk_bulge_halo_lmax12.coef", "ferrers_bar_lmax16.coef"]
EPS_BRANCH = 6.0554544523933429e-06   # cbrt(DBL_EPSILON)


def load(name):
    from oracle.pyoracle import OraclePotential
    from agama_amd.workloads import read_coef_file
    path = os.path.join(GOLDEN, name)
    return OraclePotential(path), read_coef_file(path)


def synthesis(C, k, costh, phi, mfilter=None):
    """sum over the file's (l, m) columns of C[:, k] times the normalised harmonic at (theta, phi)"""
    from scipy.special import lpmv
    out = np.zeros(np.broadcast(costh, phi).shape)
    for c in range(C.shape[0]):
        if not np.any(C[c]):
            continue
        l = math.isqrt(c)
        m = c - l * (l + 1)
        if mfilter is not None and not mfilter(m):
            continue
        am = abs(m)
        norm = math.sqrt((2 * l + 1) * math.factorial(l - am) / math.factorial(l + am))
        trig = 1.0 if m == 0 else math.sqrt(2) * (np.cos(m * phi) if m > 0 else np.sin(am * phi))
        out = out + C[c, k] * norm * lpmv(am, l, costh) * trig
    return out


def unit_vectors(n, seed):
    g = np.random.default_rng(seed).normal(size=(n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


@pytest.mark.parametrize("name", FIXTURES)
def test_node_lattice_equals_sh_synthesis(name):
    orc, (radii, Phi, dPhi) = load(name)
    tau = orc.tables(1)
    costh, sinth = 2 * tau / (1 + tau * tau), (1 - tau * tau) / (1 + tau * tau)
    worst_p = worst_d = 0.0
    for k in range(1, len(radii) - 1):        # the end nodes belong to the power-law branches
        r = radii[k]
        for ph in (0.0, 0.37, 1.9, -2.6):
            pts = np.stack([r * sinth * np.cos(ph), r * sinth * np.sin(ph), r * costh], 1)
            p, g = orc.eval(pts)
            s = synthesis(Phi, k, costh, ph)
            ds = synthesis(dPhi, k, costh, ph)
            dr = np.sum(pts * g, axis=1) / r
            worst_p = max(worst_p, np.max(np.abs(p - s) / np.abs(s)))
            worst_d = max(worst_d, np.max(np.abs(dr - ds)) / np.max(np.abs(ds)))
    print(f"{name}: lattice Phi {worst_p:.1e}, dPhi/dr {worst_d:.1e}")
    assert worst_p <= 1e-14
    assert worst_d <= 2e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_gradient_equals_finite_differences(name):
    orc, (radii, _, _) = load(name)
    rmin, rmax = radii[0], radii[-1]
    gen = np.random.default_rng(11)
    r = np.exp(gen.uniform(np.log(rmin * 1e-2), np.log(rmax * 1e2), 3000))
    # stencils of +-2h (h = 1e-3 r) that straddle a branch radius are excluded
    far = (np.abs(np.log(r / (rmin * (1 + EPS_BRANCH)))) > 0.01) & (np.abs(np.log(r / (rmax * (1 - EPS_BRANCH)))) > 0.01)
    pts = unit_vectors(3000, 12) * r[:, None]
    pts, r = pts[far], r[far]
    pot, grad = orc.eval(pts)
    fd = np.empty_like(grad)
    h = 1e-3 * r
    for c in range(3):
        def f(s):
            q = pts.copy()
            q[:, c] += s * h
            return orc.eval(q)[0]
        fd[:, c] = (f(-2) - 8 * f(-1) + 8 * f(1) - f(2)) / (12 * h)
    gnorm = np.linalg.norm(grad, axis=1)
    err = np.linalg.norm(fd - grad, axis=1) / gnorm
    # round-off of the difference quotient: ~1.5 eps |Phi| / h, large where the gradient
    # vanishes (a constant-density core inside r_min)
    noise = 1.5 * 2.2e-16 * np.abs(pot) / h / gnorm
    inside = (r > rmin) & (r < rmax)
    print(f"{name}: FD inside median {np.median(err[inside]):.1e} max {np.max(err[inside]):.1e}; "
          f"outside median {np.median(err[~inside]):.1e} max {np.max(err[~inside]):.1e}; "
          f"worst err / (1e-7 + 10 x round-off) {np.max(err / (1e-7 + 10 * noise)):.3f}")
    assert np.median(err[inside]) <= 1e-10 and np.max(err[inside]) <= 1e-7
    assert np.all(err <= 1e-7 + 10 * noise)


@pytest.mark.parametrize("name", FIXTURES)
def test_continuity_across_branch_radii(name):
    orc, (radii, Phi, _) = load(name)
    d = unit_vectors(200, 13)
    for rb, inner in ((radii[0] * (1 + EPS_BRANCH), True), (radii[-1] * (1 - EPS_BRANCH), False)):
        pa, ga = orc.eval(d * rb * (1 - 1e-12))
        pb, gb = orc.eval(d * rb * (1 + 1e-12))
        jump = pb - pa
        gjump = np.linalg.norm(gb - ga, axis=1) / np.linalg.norm(ga, axis=1)
        if name.startswith("ferrers") and not inner:
            # the reference's outer extrapolation keeps no m != 0 term for this compact model:
            # the jump is minus their synthesis at r_max (negligible l >= 4 terms aside)
            costh = d[:, 2]
            phi = np.arctan2(d[:, 1], d[:, 0])
            dropped = synthesis(Phi, len(radii) - 1, costh, phi, mfilter=lambda m: m != 0)
            print(f"{name} r_max: jump {np.max(np.abs(jump)):.2e}, m != 0 synthesis {np.max(np.abs(dropped)):.2e}")
            assert np.max(np.abs(dropped)) > 1e3 * 1e-9 * np.max(np.abs(pa))    # the effect is real
            assert np.max(np.abs(jump + dropped)) <= 1e-3 * np.max(np.abs(dropped))
            continue
        rel = np.max(np.abs(jump) / np.abs(pa))
        print(f"{name} {'r_min' if inner else 'r_max'}: Phi jump {rel:.1e}, grad jump {np.max(gjump):.1e}")
        assert rel <= 1e-9
        assert np.max(gjump) <= 2e-6


@pytest.mark.parametrize("name,mass,bar", [("dehnen_triax_lmax8.coef", 1.0, 1e-5),
                                           ("ferrers_bar_lmax16.coef", 0.02, 1e-4)])
def test_far_field_monopole(name, mass, bar):
    orc, (radii, _, _) = load(name)
    d = unit_vectors(100, 14)
    for R in (1e3 * radii[-1], 1e5 * radii[-1]):
        p, g = orc.eval(d * R)
        e_phi = np.max(np.abs(p * R / -mass - 1))
        e_g = np.max(np.abs(np.linalg.norm(g, axis=1) * R * R / mass - 1))
        print(f"{name} r={R:.0e}: Phi r / -M - 1 {e_phi:.1e}, |grad| r^2 / M - 1 {e_g:.1e}")
        assert e_phi <= bar and e_g <= bar
        assert np.all(np.sum(g * d, axis=1) > 0)       # attraction towards the centre


def _dehnen_density(x, y, z, M=1.0, a=1.0, gamma=1.0, p=0.8, q=0.6):
    m = np.sqrt(x * x + (y / p) ** 2 + (z / q) ** 2)
    return M * a * (3 - gamma) / (4 * np.pi * p * q) * m ** (-gamma) * (a + m) ** (gamma - 4)


def test_poisson_route_equals_independent_quadrature(tmp_path):
    from numpy.polynomial.legendre import leggauss
    from scipy.special import lpmv
    from agama_amd import lib, _native
    from agama_amd.workloads import read_coef_file
    path = str(tmp_path / "dehnen.coef")
    assert lib.agx_set_construction(1) == 0          # the host solver (no GPU here)
    try:
        rc = lib.agx_build_coef_file(b"type=Multipole, density=Dehnen, gamma=1, axisRatioY=0.8, axisRatioZ=0.6, "
                                     b"lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3", path.encode())
    finally:
        lib.agx_set_construction(-1)
    assert rc == 0, _native.last_error()
    radii, Phi, dPhi = read_coef_file(path)
    terms = [(l, m) for l in range(0, 9, 2) for m in range(0, l + 1, 2)]
    # rho_lm(s) = (1/4pi) int rho B_lm dOmega, B_lm the synthesis basis above
    ct, wt = leggauss(96)
    ph = (np.arange(96) + 0.5) * 2 * np.pi / 96
    CT, PH = np.meshgrid(ct, ph, indexing="ij")
    ST = np.sqrt(1 - CT ** 2)
    W = np.outer(wt, np.full(96, 2 * np.pi / 96)) / (4 * np.pi)
    B = np.array([math.sqrt((2 * l + 1) * math.factorial(l - m) / math.factorial(l + m)) * lpmv(m, l, CT) *
                  (1.0 if m == 0 else math.sqrt(2) * np.cos(m * PH)) * W for l, m in terms]).reshape(len(terms), -1)
    dx, dy, dz = (ST * np.cos(PH)).ravel(), (ST * np.sin(PH)).ravel(), CT.ravel()
    # composite Gauss-Legendre (20 nodes) in ln s: 12 pieces per grid segment, 60 decades' worth of
    # log pieces below r_0 and above r_max
    gx, gw = leggauss(20)
    edges = [np.logspace(np.log10(radii[0]) - 12, np.log10(radii[0]), 61)]
    for k in range(len(radii) - 1):
        edges.append(np.exp(np.linspace(np.log(radii[k]), np.log(radii[k + 1]), 13))[1:])
    edges.append(np.logspace(np.log10(radii[-1]), np.log10(radii[-1]) + 12, 61)[1:])
    edges = np.concatenate(edges)
    lo, hi = np.log(edges[:-1]), np.log(edges[1:])
    u = 0.5 * (hi - lo)[:, None] * (gx[None, :] + 1) + lo[:, None]
    s = np.exp(u).ravel()
    ws = (0.5 * (hi - lo)[:, None] * gw[None, :]).ravel() * s
    rho = np.empty((len(terms), len(s)))
    for i0 in range(0, len(s), 512):
        ss = s[i0:i0 + 512, None]
        rho[:, i0:i0 + 512] = B @ _dehnen_density(ss * dx, ss * dy, ss * dz).T
    nseg_below = np.searchsorted(edges, radii * (1 + 1e-12)) - 1
    worst = 0.0
    for i, (l, m) in enumerate(terms):
        c = l * (l + 1) + m
        seg_in = (rho[i] * s ** (l + 2) * ws).reshape(-1, 20).sum(1)
        seg_out = (rho[i] * s ** (1.0 - l) * ws).reshape(-1, 20).sum(1)
        Iin = np.concatenate([[0], np.cumsum(seg_in)])[nseg_below]
        Iout = np.concatenate([np.cumsum(seg_out[::-1])[::-1], [0]])[nseg_below]
        ex = -4 * np.pi / (2 * l + 1) * (radii ** (-l - 1) * Iin + radii ** l * Iout)
        exd = -4 * np.pi / (2 * l + 1) * (-(l + 1) * radii ** (-l - 2) * Iin + l * radii ** (l - 1) * Iout)
        e = max(np.max(np.abs(Phi[c] - ex) / np.abs(Phi[0])), np.max(np.abs(dPhi[c] - exd) / np.abs(dPhi[0])))
        print(f"(l, m) = ({l}, {m}): |Phi_lm / Phi_00| up to {np.max(np.abs(ex / Phi[0])):.1e}, "
              f"difference {e:.1e} of Phi_00")
        worst = max(worst, e)
    assert worst <= 1e-8


def synthetic_general_coefs():
    """coefficients with every (l, m) of lmax = 2 non-zero (odd l, sine terms: no symmetry, so
    the general synthesis path rather than the lmax = 2 fast path): Phi_00 = 3 f, the others
    a_lm g, f = -1/sqrt(r^2+1), g = -r^2/(r^2+1)^(5/2), derivatives in closed form"""
    r = np.logspace(-2, 2, 30)
    f, df = -1 / np.sqrt(r * r + 1), r / (r * r + 1) ** 1.5
    g = -r * r / (r * r + 1) ** 2.5
    dg = (-2 * r * (r * r + 1) + 5 * r ** 3) / (r * r + 1) ** 3.5
    a = np.random.default_rng(3).uniform(-0.2, 0.2, 9)
    a[0] = 3.0
    P = np.array([a[c] * (f if c == 0 else g) for c in range(9)])
    dP = np.array([a[c] * (df if c == 0 else dg) for c in range(9)])
    return r, P, dP


@pytest.mark.parametrize("case", ["plummer_lmax4.coef", "dehnen_triax_lmax2.coef", "general_lmax2"])
def test_interp1d_pins(case):
    """the reference's 1-d variant for lmax <= 2 (MultipoleInterp1d, src/potential_multipole.cpp:
    1691-1815): one quintic spline in ln r per (l, m) term, so at every grid radius the potential
    in ANY direction equals the direct synthesis of the file's coefficients (and dPhi/dr that of
    the #dPhi/dr block); gradient = finite differences; continuity across the branch radii; the
    Plummer expansion (l = 0 only) equals -M/sqrt(r^2+a^2) between the nodes"""
    from oracle.pyoracle import OraclePotential
    from agama_amd.workloads import read_coef_file
    radii, P, dP = synthetic_general_coefs() if case == "general_lmax2" else read_coef_file(os.path.join(GOLDEN, case))
    orc = OraclePotential(coefs=(radii, P, dP))
    d = unit_vectors(64, 5)
    costh, phi = d[:, 2], np.arctan2(d[:, 1], d[:, 0])
    worst_p = worst_d = 0.0
    for k in range(1, len(radii) - 1):
        p, g = orc.eval(d * radii[k])
        s, ds = synthesis(P, k, costh, phi), synthesis(dP, k, costh, phi)
        worst_p = max(worst_p, np.max(np.abs(p - s) / np.abs(s)))
        worst_d = max(worst_d, np.max(np.abs(np.sum(d * g, axis=1) - ds)) / np.max(np.abs(ds)))
    # finite differences over and beyond the grid
    gen = np.random.default_rng(11)
    r = np.exp(gen.uniform(np.log(radii[0] * 1e-2), np.log(radii[-1] * 1e2), 2000))
    far = (np.abs(np.log(r / (radii[0] * (1 + EPS_BRANCH)))) > 0.01) & \
        (np.abs(np.log(r / (radii[-1] * (1 - EPS_BRANCH)))) > 0.01)
    pts = (unit_vectors(2000, 12) * r[:, None])[far]
    r = r[far]
    pot, grad = orc.eval(pts)
    h = 1e-3 * r
    fd = np.empty_like(grad)
    for c in range(3):
        def f(s):
            q = pts.copy()
            q[:, c] += s * h
            return orc.eval(q)[0]
        fd[:, c] = (f(-2) - 8 * f(-1) + 8 * f(1) - f(2)) / (12 * h)
    gnorm = np.linalg.norm(grad, axis=1)
    err = np.linalg.norm(fd - grad, axis=1) / gnorm
    noise = 1.5 * 2.2e-16 * np.abs(pot) / h / gnorm
    inside = (r > radii[0]) & (r < radii[-1])
    print(f"{case}: nodes Phi {worst_p:.1e}, dPhi/dr {worst_d:.1e}; FD inside median {np.median(err[inside]):.1e}")
    assert worst_p <= 1e-14 and worst_d <= 1e-11
    assert np.median(err[inside]) <= 1e-10 and np.all(err <= 1e-7 + 10 * noise)
    for rb in (radii[0] * (1 + EPS_BRANCH), radii[-1] * (1 - EPS_BRANCH)):
        pa, ga = orc.eval(d * rb * (1 - 1e-12))
        pb, gb = orc.eval(d * rb * (1 + 1e-12))
        assert np.max(np.abs(pb - pa) / np.abs(pa)) <= 1e-9
        assert np.max(np.linalg.norm(gb - ga, axis=1) / np.linalg.norm(ga, axis=1)) <= 1e-5
    if case.startswith("plummer"):
        r = np.exp(np.random.default_rng(4).uniform(np.log(radii[0] * 1.01), np.log(radii[-1] * 0.99), 2000))
        p, _ = orc.eval(unit_vectors(2000, 6) * r[:, None])
        assert np.max(np.abs(p / (-1 / np.sqrt(r * r + 1)) - 1)) <= 1e-7
