"""Writes the coefficient-file fixtures used by the parity tests.

The coefficients are produced by agama_amd's host-side Poisson construction (no GPU needed,
agx_build_coef_file) from the density profiles of the workloads; both the GPU path and the
CPU oracle then load the same file, which is the parity boundary for evaluation and
integration.  Run: python tests/golden/make_fixtures.py"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from agama_amd import lib, _native  # noqa: E402

FIXTURES = {
    # triaxial Dehnen gamma=1, axis ratios 1:0.8:0.6, lmax=8, 50 radial nodes in [1e-3, 1e3]
    "dehnen_triax_lmax8.coef": "type=Multipole, density=Dehnen, gamma=1, mass=1, scaleRadius=1, "
                               "axisRatioY=0.8, axisRatioZ=0.6, lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3",
    # the same density expanded with lmax=2, 30 nodes: the reference evaluates it with 1-d splines
    # per (l, m) term and the lmax=2 fast path of the spherical-harmonic synthesis
    "dehnen_triax_lmax2.coef": "type=Multipole, density=Dehnen, gamma=1, mass=1, scaleRadius=1, "
                               "axisRatioY=0.8, axisRatioZ=0.6, lmax=2, gridSizeR=30, rmin=1e-3, rmax=1e3",
    # spherical Plummer (closed-form potential -M/sqrt(r^2+a^2)) expanded with lmax=4
    "plummer_lmax4.coef": "type=Multipole, density=Plummer, mass=1, scaleRadius=1, lmax=4, "
                          "gridSizeR=40, rmin=1e-2, rmax=1e2, symmetry=Triaxial",
    # axisymmetric composite: exponential disk + Hernquist bulge + NFW halo, lmax=12, 80 nodes
    "disk_bulge_halo_lmax12.coef": "type=Multipole, density=Disk+Dehnen+NFW, lmax=12, gridSizeR=80, "
                                   "rmin=1e-2, rmax=1e3, 0.mass=0.05, 0.scaleRadius=3, 0.scaleHeight=0.3, "
                                   "1.mass=0.01, 1.scaleRadius=0.5, 1.gamma=1, 2.mass=1, 2.scaleRadius=20",
    # Ferrers bar a=3, b=1, c=0.8 (axis ratios 1/3, 0.8/3), M=0.02, lmax=16 (even l and m for
    # its triaxial symmetry), 25 radial nodes in [0.01, 100]; used with an NFW halo in a composite
    "ferrers_bar_lmax16.coef": "type=Multipole, density=Ferrers, mass=0.02, scaleRadius=3, "
                               "axisRatioY=0.3333333333333333, axisRatioZ=0.26666666666666666, "
                               "lmax=16, gridSizeR=25, rmin=0.01, rmax=100",
}


def main(only=None):
    for name, params in FIXTURES.items():
        if only and name not in only:
            continue
        path = os.path.join(HERE, name)
        rc = lib.agx_build_coef_file(params.encode(), path.encode())
        if rc:
            raise RuntimeError(f"{name}: {_native.last_error()}")
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
