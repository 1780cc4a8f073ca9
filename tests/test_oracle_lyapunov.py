"""Pins of the oracle's Lyapunov-exponent chain (oracle/agama_oracle.cpp and oracle/pyoracle.py,
restating src/orbit_variational.cpp:121-199, 360-448, src/potential_utils.cpp:264-289, 608-619,
src/math_core.cpp:866-1040 and src/potential_base.cpp:257-283, 529-575):
  - T_circ(E): the circular orbit of the azimuthally averaged potential found independently
    (brentq on the average of oracle evaluations at the 9 azimuths);
  - the fit: synthetic bins drawn from the model with known (lambda, t_chaos) are recovered;
    linear growth gives (0, inf); fewer than 5 bins give NaN;
  - whole orbits: orbits whose deviation vector grows exponentially (beyond the linear growth
    of regular orbits) get a positive exponent, the others (0, inf)."""
import math
import os

import numpy as np
import pytest
from scipy.optimize import brentq

from tests.conftest import GOLDEN
from oracle.pyoracle import OraclePotential, lyapunov_from_bins

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def orc():
    return OraclePotential(COEF)


def _avg(orc, R):
    """potential and dPhi/dR averaged over 2*8+1 azimuths of a y-symmetric expansion"""
    phis = np.pi / 8.5 * np.arange(9)
    pts = np.stack([R * np.cos(phis), R * np.sin(phis), np.zeros(9)], 1)
    p, g = orc.eval(pts)
    dR = g[:, 0] * np.cos(phis) + g[:, 1] * np.sin(phis)
    w = np.where(np.arange(9) > 0, 2.0, 1.0)
    return np.sum(p * w) / 17, np.sum(dR * w) / 17


def test_tcirc_matches_independent_root(orc):
    for E in (-0.9, -0.6, -0.3, -0.1, -0.02):
        R = brentq(lambda R: (lambda a: a[0] + 0.5 * R * a[1] - E)(_avg(orc, R)), 1e-3, 1e3, xtol=1e-14, rtol=1e-14)
        vc = math.sqrt(R * _avg(orc, R)[1])
        ref = 2 * math.pi * R / vc
        got = orc.tcirc(E)
        assert abs(got / ref - 1) < 1e-8, (E, got, ref)


def _model(t, lam, A, tch):
    from oracle.pyoracle import _logexpm1overx as F
    return A + np.log(t) + np.where(t < tch, 0, F(lam * t) - F(lam * tch))


def test_fit_recovers_known_parameters():
    t = np.linspace(5, 500, 60)
    lam, A, tch, period = 0.05, -1.3, 180.0, 7.0
    out = lyapunov_from_bins(t, _model(t, lam, A, tch), 505.0, period)
    assert abs(out[0] / (lam * period) - 1) < 1e-6 and abs(out[1] / (tch / period) - 1) < 1e-6
    # linear growth only: no chaos detected
    out = lyapunov_from_bins(t, -1.3 + np.log(t), 505.0, period)
    assert out[0] == 0 and out[1] == np.inf
    # too few bins / nothing recorded
    assert np.all(np.isnan(lyapunov_from_bins(t[:4], _model(t[:4], lam, A, tch), 505.0, period)))
    assert np.all(np.isnan(lyapunov_from_bins([], [], np.nan, period)))


def test_orbits_regular_and_chaotic(orc):
    from agama_amd import workloads
    ic, T = workloads.triaxial_dehnen_orbits(16, lambda p: orc.eval(p)[0], workloads.read_coef_file(COEF),
                                             seed=11, periods=100)
    r = orc.orbits_lyap(ic, T)
    lyap, nb, period = r["lyap"], r["nbins"], r["period"]
    assert np.all(period > 0) and np.all(nb >= 5)
    chaotic = lyap[:, 0] > 0
    assert 0 < chaotic.sum() < len(ic)           # this sample has both kinds
    assert np.all(lyap[~chaotic, 1] == np.inf)
    for i in range(len(ic)):
        # the bins' growth beyond the linear (ln t) part separates the two kinds
        t, lnw = r["bins"][i, 0, :nb[i]], r["bins"][i, 1, :nb[i]]
        excess = (lnw[-1] - np.log(t[-1])) - (lnw[0] - np.log(t[0]))
        if chaotic[i]:
            assert excess > math.log(2), i
        else:
            assert excess < 5, i
    # der=True follows six vectors; the fiducial combination grows like the single vector
    r6 = orc.orbits_lyap(ic[:4], T[:4], der=True)
    assert np.array_equal(r6["period"], period[:4])
    assert np.array_equal(r6["lyap"][:, 0] > 0, chaotic[:4])


def _lib_fit(t, lnw, max_time, period):
    from agama_amd._native import lib
    t = np.ascontiguousarray(t, dtype=np.float64)
    lnw = np.ascontiguousarray(lnw, dtype=np.float64)
    out = np.empty(2)
    assert lib.agx_lyapunov_fit(t.ctypes.data, lnw.ctypes.data, len(t), float(max_time), float(period),
                                out.ctypes.data) == 0
    return out


def test_library_fit_matches_oracle_fit(orc):
    """the library's host-side fit (its own Levenberg-Marquardt with the reference's stopping
    rule) against the oracle's MINPACK fit on identical bins: synthetic ones and those of real
    orbits (host code only; no device call)"""
    t = np.linspace(5, 500, 60)
    for lam, tch in ((0.05, 180.0), (0.02, 0.0), (0.2, 350.0), (0.0, 0.0)):
        lnw = _model(t, lam, -1.3, tch) if lam > 0 else -1.3 + np.log(t)
        a, b = _lib_fit(t, lnw, 505.0, 7.0), lyapunov_from_bins(t, lnw, 505.0, 7.0)
        assert (a[0] > 0) == (b[0] > 0)
        if b[0] > 0:
            assert abs(a[0] / b[0] - 1) < 1e-5 and abs(a[1] - b[1]) <= 1e-5 * max(b[1], 1)
    from agama_amd import workloads
    ic, T = workloads.triaxial_dehnen_orbits(16, lambda p: orc.eval(p)[0], workloads.read_coef_file(COEF),
                                             seed=11, periods=100)
    r = orc.orbits_lyap(ic, T)
    rel = []
    for i in range(len(ic)):
        nb = r["nbins"][i]
        a = _lib_fit(r["bins"][i, 0, :nb], r["bins"][i, 1, :nb], r["last"][i], r["period"][i])
        b = r["lyap"][i]
        assert (a[0] > 0) == (b[0] > 0), i
        if b[0] > 0:
            rel.append(abs(a[0] / b[0] - 1))
    print(f"library vs oracle fit on orbit bins: rel diff median {np.median(rel):.2e} max {max(rel):.2e}")
    # with the onset before the first bin the model has a flat direction (tch trades against A),
    # where the two solvers may stop at slightly different points; elsewhere they agree closely
    assert np.median(rel) < 1e-5 and max(rel) < 2e-2
    assert np.all(np.isnan(_lib_fit([], [], np.nan, 1.0)))
