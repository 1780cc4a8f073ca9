"""The Poisson solver in the device's form, checked on the CPU: the density at the quadrature
points, the forward spherical-harmonic transform as one matrix product (instead of the
reference's Fourier-then-Legendre passes) with the near-zero cut, and the radial integrals
(poissonStagesHost, the same operations as agx_construct.hip) against the construction in the
reference's order (multipoleFromDensity, src/potential_multipole.cpp:1013-1114, 238-298).
Bar: every coefficient to 2e-15 of its radius's largest one, identical zero pattern (symmetry
and the 1e-10 cut), for triaxial, axisymmetric, composite and spherical sources."""
import numpy as np
import pytest

CASES = [
    "type=Multipole, density=Dehnen, gamma=1, axisRatioY=0.8, axisRatioZ=0.6, lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3",
    "type=Multipole, density=Ferrers, mass=1, scaleRadius=2, axisRatioY=0.7, axisRatioZ=0.4, lmax=10, gridSizeR=40, "
    "rmin=1e-2, rmax=1e2",
    "type=Multipole, density=Disk+Dehnen, 0.mass=0.05, 0.scaleRadius=3, 0.scaleHeight=0.3, 1.mass=1, 1.scaleRadius=1, "
    "1.gamma=1.5, lmax=12, gridSizeR=60, rmin=1e-3, rmax=1e3",
    "type=Multipole, density=Plummer, mass=2, scaleRadius=0.5, lmax=4, gridSizeR=30, rmin=1e-3, rmax=1e2",
    "type=Multipole, density=NFW, mass=3, scaleRadius=5, lmax=6, gridSizeR=30, rmin=1e-2, rmax=1e3",
]


def build(params, mode, path):
    from agama_amd._native import lib
    from agama_amd import workloads
    assert lib.agx_set_construction(mode) == 0
    try:
        rc = lib.agx_build_coef_file(params.encode(), str(path).encode())
    finally:
        lib.agx_set_construction(-1)
    assert rc == 0, lib.agx_last_error()
    return workloads.read_coef_file(str(path))


@pytest.mark.parametrize("params", CASES)
def test_staged_equals_reference_order(params, tmp_path):
    r1, p1, d1 = build(params, 1, tmp_path / "host.coef")
    r2, p2, d2 = build(params, 2, tmp_path / "staged.coef")
    assert np.array_equal(r1, r2) and p1.shape == p2.shape
    assert np.array_equal(p1 == 0, p2 == 0) and np.array_equal(d1 == 0, d2 == 0)
    assert np.max(np.abs(p1 - p2) / np.abs(p1).max(axis=0)) <= 2e-15
    assert np.max(np.abs(d1 - d2) / np.abs(d1).max(axis=0)) <= 2e-15


def test_construction_mode_is_checked():
    from agama_amd._native import lib
    assert lib.agx_set_construction(7) != 0
