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


@pytest.mark.parametrize("degree", [0, 1])
def test_cylindrical_targets_closed_forms(degree):
    """TargetDensityCylindrical<0/1> (galaxymodel_densitygrid.cpp:486-545): a circular orbit in
    the equatorial plane puts its whole weight at z = 0 and R = r0 (one top-hat cell; the two
    bilinear R-nodes around r0 at the z = 0 row, linearly), with the m > 0 terms averaging to
    zero; a radial orbit on the long axis has cos(2 phi) = 1, so every m = 2 value is twice the
    m = 0 value of the same cell or R > 0 node; the m = 0 basis partitions unity."""
    from agama_amd import workloads
    gridR = np.array([0.2, 0.5, 1.0, 2.0, 5.0])
    gridz = np.array([0.1, 0.4, 1.0, 3.0])
    nR, nz, mmax = 5, 4, 2
    orc = OraclePotential(os.path.join(GOLDEN, "disk_bulge_halo_lmax12.coef"))
    r0 = 0.7
    _, g = orc.eval(np.array([[r0, 0.0, 0.0]]))
    vc = math.sqrt(r0 * abs(g[0, 0]))
    period = 2 * math.pi * r0 / vc
    out = orc.orbits_target_cyl(np.array([[r0, 0, 0, 0, vc, 0]]), 5 * period, gridR, gridz, mmax, degree,
                                accuracy=1e-10)[0]
    assert len(out) == (nz + degree) * (nR * 2 + degree)
    expected = np.zeros(len(out))
    offR = (r0 - 0.5) / (1.0 - 0.5)
    if degree == 0:
        expected[2] = 1.0                      # indR = 2, indz = 0, m = 0
    else:
        expected[2] = 1 - offR                 # node R = 0.5 (index 2 with R = 0 at 0), z-row 0
        expected[3] = offR
    # the m = 2 terms carry 2 cos(2 phi): their time average over whole turns vanishes to the
    # sampling of ten points per step (a few 1e-6)
    assert np.max(np.abs(out - expected)) < 1e-5, (out, expected)
    # radial orbit on the long axis of the triaxial model
    coef = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    tri = OraclePotential(coef)
    rad = tri.orbits_target_cyl(np.array([[3.0, 0, 0, 0, 0, 0]]), 40.0, gridR, gridz, mmax, degree)[0]
    if degree == 0:
        m0, m2 = rad[:nR * nz], rad[nR * nz:]
        assert np.allclose(m2, 2 * m0, rtol=2e-6, atol=1e-9)
    else:
        m0 = rad[:(nR + 1) * (nz + 1)].reshape(nz + 1, nR + 1)[:, 1:]      # R > 0 nodes
        m2 = rad[(nR + 1) * (nz + 1):].reshape(nz + 1, nR)
        assert np.allclose(m2, 2 * m0, rtol=2e-6, atol=1e-9)
    # partition of unity of the m = 0 terms for orbits inside the grid
    ic, T = workloads.triaxial_dehnen_orbits(12, lambda p: tri.eval(p)[0], workloads.read_coef_file(coef),
                                             seed=52, periods=4)
    r = np.linalg.norm(ic[:, :3], axis=1)
    sel = r < 0.5
    big = np.concatenate([[0.0], np.logspace(-3, 3, 30)])
    res = tri.orbits_target_cyl(ic[sel], T[sel], big, big, mmax, degree)
    m0 = res[:, :(30 + degree) * (30 + degree)]
    assert np.max(np.abs(m0.sum(axis=1) - 1)) < 2e-6


def _scipy_bspline_basis(degree, grid, x):
    """the clamped B-spline basis of scipy (end knots repeated degree times) at points x"""
    from scipy.interpolate import BSpline
    knots = np.concatenate([[grid[0]] * degree, grid, [grid[-1]] * degree])
    return BSpline.design_matrix(x, knots, degree).toarray()


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_kinem_shell_bspline_basis_matches_scipy(degree):
    """BsplineInterpolator1d<N> (math_spline.cpp:35-82, 1877-1886): de Boor's recurrence with
    node positions clamped at the grid ends is the clamped B-spline basis; it partitions unity
    inside the grid and is zero outside it"""
    from oracle.pyoracle import bspline_values
    grid = np.array([0.0, 0.1, 0.25, 0.3, 0.7, 1.5, 2.0, 4.0])
    rng = np.random.default_rng(71)
    x = np.concatenate([rng.uniform(0, 4.0, 300), grid[:-1]])
    B = bspline_values(degree, grid, x)
    assert B.shape == (len(x), len(grid) + degree - 1)
    assert np.max(np.abs(B - _scipy_bspline_basis(degree, grid, x))) < 1e-14
    assert np.max(np.abs(B.sum(axis=1) - 1)) < 1e-14
    assert np.all(bspline_values(degree, grid, [-0.1, 4.1]) == 0)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_kinem_shell_circular_and_radial_orbits(degree):
    """TargetKinemShell<N> (galaxymodel_losvd.cpp:568-576): a circular orbit in the equatorial
    plane of an axisymmetric potential has v_r = 0 and v_t = v_c at one radius, so the v_t^2
    values are v_c^2 times the basis at r0 and the v_r^2 values vanish, in the inertial frame and
    in a rotating one (the target sees velocities rotated, not corotating ones); an orbit released
    from rest on the long axis of the triaxial model has v_t = 0, and its v_r^2 values are
    positive where the orbit goes"""
    orc = OraclePotential(os.path.join(GOLDEN, "disk_bulge_halo_lmax12.coef"))
    r0 = 0.7
    _, g = orc.eval(np.array([[r0, 0.0, 0.0]]))
    vc2 = r0 * abs(g[0, 0])
    period = 2 * math.pi * r0 / math.sqrt(vc2)
    gridr = np.array([0.0, 0.2, 0.5, 1.0, 2.0, 5.0])
    nb = len(gridr) + degree - 1
    expected = np.concatenate([np.zeros(nb), vc2 * _scipy_bspline_basis(degree, gridr, np.array([r0]))[0]])
    for Omega in (0.0, 0.4):
        out = orc.orbits_target_kinem_shell(np.array([[r0, 0, 0, 0, math.sqrt(vc2), 0]]), 5 * period, gridr,
                                            degree, accuracy=1e-10, Omega=Omega)[0]
        assert len(out) == 2 * nb
        assert np.max(np.abs(out - expected)) < 1e-5 * vc2, (Omega, out, expected)
    tri = OraclePotential(os.path.join(GOLDEN, "dehnen_triax_lmax8.coef"))
    rad = tri.orbits_target_kinem_shell(np.array([[3.0, 0, 0, 0, 0, 0]]), 40.0, gridr, degree)[0]
    vr2, vt2 = rad[:nb], rad[nb:]
    assert np.max(np.abs(vt2)) < 1e-8 * np.max(vr2)
    assert np.all(vr2 >= 0) and np.sum(vr2 > 0) >= 4
    # zero-length orbits record nothing
    assert np.all(tri.orbits_target_kinem_shell(np.array([[3.0, 0, 0, 0, 0, 0]]), 0.0, gridr, degree) == 0)


@pytest.mark.parametrize("degree", [0, 1])
def test_classic_targets_closed_forms(degree):
    """TargetDensityClassic<0/1> (galaxymodel_densitygrid.cpp:146-239, corner numbering :59-90).
    A circular orbit in the equatorial plane of an axisymmetric potential, folded into the first
    quadrant, moves uniformly in the angle psi = atan(|y|/|x|) through pane 0 (psi <= pi/4, strips
    of pi/(4S) in atan(Y/X)) and pane 1 (psi > pi/4, strips in atan(X/Y)), always in the first row
    of the other direction: each of the 2S top-hat cells holds 1/(2S) of the time; the trilinear
    nodes along that arc are hat functions in psi, 1/(2S) each and 1/(4S) at the two ends (the
    node at the diagonal shared by both panes through the adjacent-pane numbering), split
    linearly between the two shells around r0.  Squeezed axes: on a radial orbit along x (or z)
    the shells sit at radius r * axisX (axisZ), so the result equals that of a spherical grid
    scaled by that axis.  The basis partitions unity inside the grid."""
    from agama_amd import workloads
    from oracle.pyoracle import classic_target_size
    S = 3
    gridr = np.array([0.0, 0.2, 0.5, 1.0, 2.0])
    orc = OraclePotential(os.path.join(GOLDEN, "disk_bulge_halo_lmax12.coef"))
    r0 = 0.7
    _, g = orc.eval(np.array([[r0, 0.0, 0.0]]))
    vc = math.sqrt(r0 * abs(g[0, 0]))
    period = 2 * math.pi * r0 / vc
    out = orc.orbits_target_classic(np.array([[r0, 0, 0, 0, vc, 0]]), 5 * period, gridr, degree, S,
                                    accuracy=1e-10)[0]
    assert len(out) == classic_target_size(degree, S, gridr) == (3 * S * (S + degree) + degree) * 4 + degree
    expected = np.zeros(len(out))
    if degree == 0:
        vps, sh = 3 * S * S, 2
        for k in range(S):
            expected[sh * vps + k] = 1 / (2 * S)                    # pane 0, strip k, first row
            expected[sh * vps + (k + S) * S] = 1 / (2 * S)          # pane 1, first strip, row k
    else:
        vps, offr = 3 * S * (S + 1) + 1, (r0 - 0.5) / (1.0 - 0.5)
        w = np.full(2 * S + 1, 1 / (2 * S))
        w[0] = w[-1] = 1 / (4 * S)
        nodes = [k for k in range(S + 1)] + [(j + S) * (S + 1) for j in range(S - 1, -1, -1)]
        for sh, frac in ((2, offr), (1, 1 - offr)):
            for node, wk in zip(nodes, w):
                expected[sh * vps + 1 + node] += frac * wk
    # the samples (ten per step) misplace at most half a sample at each of the 8 crossings of a
    # strip boundary per period: the time fractions agree to 4 / (samples per period)
    nsteps = orc.orbits(np.array([[r0, 0, 0, 0, vc, 0]]), 5 * period, trajsize=1, accuracy=1e-10)["nsteps"][0]
    bound = 4 / (10 * nsteps / 5)
    assert np.max(np.abs(out - expected)) < bound, (np.max(np.abs(out - expected)), bound)
    assert abs(out.sum() - 1) < 2e-6
    # squeezed grid on radial orbits along the x and z axes of the triaxial model
    coef = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    tri = OraclePotential(coef)
    qy, qz = 0.6, 0.8
    ax = 1 / np.cbrt(qy * qz)
    for ic, scale in (([3.0, 0, 0, 0, 0, 0], ax), ([0, 0, 2.0, 0, 0, 0], qz * ax)):
        sq = tri.orbits_target_classic(np.array([ic]), 30.0, gridr, degree, S, qy, qz)[0]
        sph = tri.orbits_target_classic(np.array([ic]), 30.0, gridr * scale, degree, S)[0]
        assert np.count_nonzero(sq) >= 4
        assert np.allclose(sq, sph, rtol=1e-6, atol=1e-9)
    # partition of unity
    ic, T = workloads.triaxial_dehnen_orbits(12, lambda p: tri.eval(p)[0], workloads.read_coef_file(coef),
                                             seed=53, periods=4)
    big = np.concatenate([[0.0], np.logspace(-3, 4, 36)])
    res = tri.orbits_target_classic(ic, T, big, degree, 4, 0.7, 0.5)
    assert np.max(np.abs(res.sum(axis=1) - 1)) < 2e-6


def test_targets_with_dprkn8_closed_forms():
    """the same target restatements fed by DPRKN8's dense output (any integrator drives a
    RuntimeFncTarget): the circular-orbit values and the partition of unity hold as with DOP853"""
    from agama_amd import workloads
    orc = OraclePotential(os.path.join(GOLDEN, "disk_bulge_halo_lmax12.coef"))
    r0 = 0.7
    _, g = orc.eval(np.array([[r0, 0.0, 0.0]]))
    vc = math.sqrt(r0 * abs(g[0, 0]))
    period = 2 * math.pi * r0 / vc
    gridr = np.array([0.2, 0.5, 1.0, 2.0])
    out = orc.orbits_target_sphharm(np.array([[r0, 0, 0, 0, vc, 0]]), 5 * period, gridr, 4, 2, accuracy=1e-10,
                                    method="dprkn8")[0]
    offr = (r0 - 0.5) / 0.5
    expected = np.zeros(len(out))
    for blk, val in enumerate([1.0, -math.sqrt(5) / 2, 9 / 8, 0.0, 0.0]):
        expected[1 + blk * 4 + 1] = val * (1 - offr)
        expected[1 + blk * 4 + 2] = val * offr
    assert np.max(np.abs(out - expected)) < 2e-6
    kin = orc.orbits_target_kinem_shell(np.array([[r0, 0, 0, 0, vc, 0]]), 5 * period, np.array([0.0, 0.5, 1.0, 2.0]),
                                        1, accuracy=1e-10, method="dprkn8")[0]
    assert np.max(np.abs(kin[:4])) < 1e-5 * vc ** 2 and abs(kin[4:].sum() - vc ** 2) < 1e-5 * vc ** 2
    coef = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    tri = OraclePotential(coef)
    ic, T = workloads.triaxial_dehnen_orbits(8, lambda p: tri.eval(p)[0], workloads.read_coef_file(coef),
                                             seed=54, periods=3)
    big = np.concatenate([[0.0], np.logspace(-3, 4, 36)])
    res = tri.orbits_target_classic(ic, T, big, 1, 3, method="dprkn8")
    assert np.max(np.abs(res.sum(axis=1) - 1)) < 2e-6


def test_target_objects_host_side():
    """agama_amd.Target (no GPU call): len() equals the oracle's value counts for every type, and
    invalid parameters raise as in the reference's constructors"""
    import agama_amd
    from oracle.pyoracle import classic_target_size, cyl_target_size, sphharm_target_size
    gr = np.concatenate([[0.0], np.logspace(-2, 2, 9)])
    gz = np.concatenate([[0.0], np.logspace(-2, 1, 6)])
    assert len(agama_amd.Target(type="DensitySphHarm", gridr=gr, lmax=6, mmax=4)) == sphharm_target_size(6, 4, gr)
    for deg, kind in ((0, "DensityCylindricalTopHat"), (1, "DensityCylindricalLinear")):
        assert len(agama_amd.Target(type=kind, gridr=gr, gridz=gz, mmax=4)) == cyl_target_size(deg, 4, gr, gz)
    for deg, kind in ((0, "DensityClassicTopHat"), (1, "DensityClassicLinear")):
        t = agama_amd.Target(type=kind, gridr=gr, stripsPerPane=3, axisRatioY=0.8, axisRatioZ=0.6)
        assert len(t) == classic_target_size(deg, 3, gr)
    for deg in range(4):
        assert len(agama_amd.Target(type="kinemshell", gridr=gr, degree=deg)) == 2 * (len(gr) + deg - 1)
    with pytest.raises(ValueError):
        agama_amd.Target(type="DensityClassicLinear", gridr=gr, stripsPerPane=0)
    with pytest.raises(ValueError):
        agama_amd.Target(type="DensityClassicTopHat", gridr=gr, axisRatioY=-1)
    with pytest.raises(ValueError):
        agama_amd.Target(type="KinemShell", gridr=gr[::-1], degree=1)
    with pytest.raises(TypeError):
        agama_amd.Target(type="DensitySphHarm", gridr=gr, lmax=2, bogus=1)
    with pytest.raises(agama_amd.AgamaAmdError):
        agama_amd.Target(type="LOSVD", gridr=gr)
