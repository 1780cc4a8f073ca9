"""Pins of the oracle's target plug-in (RuntimeFncTarget, src/galaxymodel_target.h:95-155, with
TargetDensitySphHarm, src/galaxymodel_densitygrid.cpp:356-411), against what the mathematics
fixes:
  - a circular orbit in the equatorial plane of an axisymmetric potential, integrated over whole
    periods, spends all its time at one radius and polar angle: the l=0 term splits the unit
    weight linearly between the two radial nodes around it, the l>0, m=0 terms carry the
    normalised Legendre values at the equator (2 sqrt(pi) Y_l0(pi/2): -sqrt(5)/2 for l=2, 9/8
    for l=4), and the m>0 terms average to zero;
  - the l=0 basis functions partition unity inside the grid, so for an orbit that stays inside
    them its l=0 values sum to 1, whatever the orbit;
  - zero-length orbits record nothing; backward integration records positive weights."""
import math
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN
from oracle.pyoracle import OraclePotential, sphharm_target_size


def test_circular_orbit_closed_form():
    orc = OraclePotential(os.path.join(GOLDEN, "disk_bulge_halo_lmax12.coef"))   # axisymmetric, z-symmetric
    r0 = 0.7
    _, g = orc.eval(np.array([[r0, 0.0, 0.0]]))
    vc = math.sqrt(r0 * abs(g[0, 0]))
    period = 2 * math.pi * r0 / vc
    ic = np.array([[r0, 0, 0, 0, vc, 0]])
    gridr = np.array([0.2, 0.5, 1.0, 2.0])
    lmax, mmax = 4, 2
    out = orc.orbits_target_sphharm(ic, 5 * period, gridr, lmax, mmax, accuracy=1e-10)[0]
    assert len(out) == sphharm_target_size(lmax, mmax, gridr) == 5 * 4 + 1
    nr, offr = 4, (r0 - 0.5) / (1.0 - 0.5)
    expected = np.zeros(len(out))
    # blocks of nr values after the origin: (l,m) = (0,0), (2,0), (4,0), (2,2), (4,2)
    for blk, val in enumerate([1.0, -math.sqrt(5) / 2, 9 / 8, 0.0, 0.0]):
        expected[1 + blk * nr + 1] = val * (1 - offr)
        expected[1 + blk * nr + 2] = val * offr
    assert np.max(np.abs(out - expected)) < 2e-6, (out, expected)


def test_unit_partition_and_edge_cases():
    from agama_amd import workloads
    coef = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    orc = OraclePotential(coef)
    ic, T = workloads.triaxial_dehnen_orbits(12, lambda p: orc.eval(p)[0], workloads.read_coef_file(coef),
                                             seed=51, periods=4)
    gridr = np.concatenate([[0.0], np.logspace(-2, 3, 16)])   # a node at 0 is dropped
    out = orc.orbits_target_sphharm(ic, T, gridr, 6, 6)
    nr = 16
    l0 = out[:, 0] + out[:, 1:1 + nr].sum(axis=1)
    assert np.max(np.abs(l0 - 1)) < 2e-6
    assert np.all(np.isfinite(out))
    z = orc.orbits_target_sphharm(ic[:3], 0.0, gridr, 6, 6)
    assert np.all(z == 0)
    b = orc.orbits_target_sphharm(ic[:3], -T[:3], gridr, 6, 6)
    assert np.max(np.abs(b[:, 0] + b[:, 1:1 + nr].sum(axis=1) - 1)) < 2e-6


def test_radial_orbit_on_the_x_axis_fixes_every_angular_term():
    """an orbit released from rest on the long axis of the triaxial model stays on it (the
    expansion's reflection symmetries), so every sample sits at theta = pi/2 and phi = 0 or pi:
    at each radial node the (l, m) value over the l=0 value is 2 sqrt(pi) N_lm P_l^m(0)
    (times sqrt(2) cos(m phi) = sqrt(2) for m > 0), from scipy's associated Legendre functions"""
    from scipy.special import lpmv
    coef = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    orc = OraclePotential(coef)
    gridr = np.logspace(-1.5, 1, 10)
    lmax, mmax, nr = 6, 4, 10
    out = orc.orbits_target_sphharm(np.array([[3.0, 0, 0, 0, 0, 0]]), 40.0, gridr, lmax, mmax)[0]
    base = out[1:1 + nr]
    blk = 0
    for m in range(0, mmax + 1, 2):
        for l in range(m, lmax + 1, 2):
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            ratio = 2 * math.sqrt(math.pi) * norm * lpmv(m, l, 0.0) * (math.sqrt(2) if m else 1)
            vals = out[1 + blk * nr:1 + (blk + 1) * nr]
            sel = base > 1e-3
            assert sel.sum() >= 5
            assert np.max(np.abs(vals[sel] / base[sel] - ratio)) < 2e-5 * max(1, abs(ratio)), (l, m)
            blk += 1
