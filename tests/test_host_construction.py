"""CPU tests of the product's host side: the Multipole construction (Poisson solver ->
coefficients, coefficients -> spline tables) and the C ABI surface.  No GPU needed."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from agama_amd import _native, lib
from oracle.pyoracle import OraclePotential
from tests.conftest import GOLDEN, ROOT

FIXTURES = ["dehnen_triax_lmax8.coef", "disk_bulge_halo_lmax12.coef"]


def product_tables(path, which, m=0):
    n = lib.agx_debug_tables(path.encode(), which, m, None)
    assert n >= 0, _native.last_error()
    out = np.empty(n)
    lib.agx_debug_tables(path.encode(), which, m, out.ctypes.data)
    return out


@pytest.mark.parametrize("name", FIXTURES)
def test_tables_bitwise_equal_to_oracle(name):
    """MultipoleInterp2d construction (potential_multipole.cpp:1854-1967): grids, scaling
    constants and all nine node arrays of every harmonic agree bit for bit"""
    path = os.path.join(GOLDEN, name)
    orc = OraclePotential(path)
    for which in (0, 1, 3):
        assert np.array_equal(product_tables(path, which), orc.tables(which))
    scal = orc.tables(3)
    mmax = int(scal[3])
    nonempty = 0
    for m in range(-mmax, mmax + 1):
        a, b = product_tables(path, 2, m), orc.tables(2, m)
        assert np.array_equal(a, b), m
        nonempty += a.size > 0
    assert nonempty >= 1


def read_coefs(path):
    lines = open(path).read().splitlines()
    i = lines.index("#Phi")
    nr = int([ln for ln in lines if ln.startswith("gridSizeR=")][0].split("=")[1])
    phi = np.array([[float(v) for v in ln.split()] for ln in lines[i + 2:i + 2 + nr]])
    j = lines.index("#dPhi/dr")
    dphi = np.array([[float(v) for v in ln.split()] for ln in lines[j + 2:j + 2 + nr]])
    return phi[:, 0], phi[:, 1:].T, dphi[:, 1:].T


def test_poisson_plummer_closed_form():
    """spherical Plummer: Phi_00(r) = -M/sqrt(r^2+a^2), dPhi/dr = M r/(r^2+a^2)^1.5"""
    r, P, dP = read_coefs(os.path.join(GOLDEN, "plummer_lmax4.coef"))
    assert P.shape[0] == 1                  # spherical symmetry keeps only l=0
    assert np.max(np.abs(P[0] / (-1 / np.sqrt(r ** 2 + 1)) - 1)) < 1e-9
    assert np.max(np.abs(dP[0] / (r / (r ** 2 + 1) ** 1.5) - 1)) < 1e-8


def test_poisson_hernquist_closed_form(tmp_path):
    """spherical Hernquist (Dehnen gamma=1): Phi = -M/(r+a)"""
    path = str(tmp_path / "h.coef")
    assert lib.agx_build_coef_file(b"type=Multipole, density=Dehnen, gamma=1, mass=2, scaleRadius=1.5, "
                                   b"lmax=0, gridSizeR=40, rmin=1e-3, rmax=1e3", path.encode()) == 0
    r, P, dP = read_coefs(path)
    assert np.max(np.abs(P[0] / (-2 / (r + 1.5)) - 1)) < 1e-6
    assert np.max(np.abs(dP[0] / (2 / (r + 1.5) ** 2) - 1)) < 1e-5


def test_triaxial_symmetry_pattern():
    """triaxial symmetry: only even l and even m >= 0 are non-zero (math_sphharm.h:76-114)"""
    r, P, dP = read_coefs(os.path.join(GOLDEN, "dehnen_triax_lmax8.coef"))
    assert P.shape[0] == 81
    for c in range(81):
        l = int(np.sqrt(c))
        m = c - l * (l + 1)
        nz = np.any(P[c] != 0) or np.any(dP[c] != 0)
        assert nz == (l % 2 == 0 and m >= 0 and m % 2 == 0), (l, m)


def test_coefficient_file_roundtrip(tmp_path):
    """file -> construction -> file reproduces the coefficients exactly (17 digits)"""
    src = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
    a = read_coefs(src)
    dst = str(tmp_path / "rt.coef")
    params = ("type=Multipole, density=Dehnen, gamma=1, mass=1, scaleRadius=1, axisRatioY=0.8, "
              "axisRatioZ=0.6, lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3")
    assert lib.agx_build_coef_file(params.encode(), dst.encode()) == 0
    b = read_coefs(dst)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_invalid_parameters_report_errors():
    assert lib.agx_build_coef_file(b"type=Multipole, density=Nonsense", b"/tmp/x") != 0
    assert "unknown density" in _native.last_error()
    assert lib.agx_build_coef_file(b"type=Multipole, density=Dehnen, rmin=1, rmax=0.5", b"/tmp/x") != 0
    assert lib.agx_build_coef_file(b"type=Multipole, density=Dehnen, bogus=3, rmin=1, rmax=5", b"/tmp/x") != 0
    assert "unknown parameter" in _native.last_error()
    assert lib.agx_debug_tables(b"/nonexistent.coef", 0, 0, None) == -1


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "agama_amd.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", header)) - {"if", "defined", "sizeof"}
    names = {n for n in names if n.startswith(("agx_", "agama_")) or
             n in ("iniacceleration", "inipotential", "potential_double", "potential_float")}
    names.discard("agx_acc_pter")
    assert "agx_orbit_batch" in names and "agama_evalPotential" in names and "potential_float" in names
    out = subprocess.check_output(["nm", "-D", "--defined-only", _native.LIB_PATH]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    assert set(_native.EXPORTED) <= exported


def test_orbit_group_settings_validation():
    """agx_set_orbit_groups: host-side validation only (no GPU call); the defaults are restored"""
    lib = _native.lib
    assert lib.agx_set_orbit_groups(-2, -1, 0) != 0
    assert "group-mode" in _native.last_error()
    assert lib.agx_set_orbit_groups(4, -2, 0) != 0
    assert lib.agx_set_orbit_groups(4, 100, -1) != 0
    assert lib.agx_set_orbit_groups(0, -1, 0) == 0
    assert lib.agx_set_orbit_groups(-1, -1, 65536) == 0
    import agama_amd
    with pytest.raises(agama_amd.AgamaAmdError):
        agama_amd.set_orbit_groups(stride=-3)
    agama_amd.set_orbit_groups()


def write_coef_file(path, radii, Phi, dPhi, symmetry="None"):
    """a Multipole coefficient file in the reference's text format (potential_factory.cpp:834-852,
    909-929): header, the #Phi and #dPhi/dr blocks of (l+1)^2 columns, values as repr()"""
    nc = Phi.shape[0]
    lmax = int(round(np.sqrt(nc))) - 1
    cols = "\t".join(f"l={l},m={m}" for l in range(lmax + 1) for m in range(-l, l + 1))
    lines = ["[Potential]", "type=Multipole", f"gridSizeR={len(radii)}", f"lmax={lmax}", f"symmetry={symmetry}",
             "Coefficients"]
    for tag, block in (("#Phi", Phi), ("#dPhi/dr", dPhi)):
        lines += [tag, "#radius\t" + cols]
        lines += ["\t".join(repr(float(v)) for v in [radii[k]] + list(block[:, k])) for k in range(len(radii))]
        lines.append("")
    open(path, "w").write("\n".join(lines) + "\n")


@pytest.mark.parametrize("case", ["plummer_lmax4.coef", "dehnen_triax_lmax2.coef", "general_lmax2"])
def test_interp1d_tables_bitwise_equal_to_oracle(case, tmp_path):
    """expansions with lmax <= 2 are built as the reference's MultipoleInterp1d
    (potential_multipole.cpp:1691-1738): grid, log-scaling flag and 1/Phi_00(0), and each (l, m)
    term's spline values, ln r derivatives and second derivatives agree bit for bit"""
    from tests.test_oracle_multipole import synthetic_general_coefs
    if case == "general_lmax2":
        path = str(tmp_path / "general.coef")
        write_coef_file(path, *synthetic_general_coefs())
    else:
        path = os.path.join(GOLDEN, case)
    orc = OraclePotential(path)
    assert np.array_equal(product_tables(path, 0), orc.tables(0))
    assert np.array_equal(product_tables(path, 3), orc.tables(3))
    lmax = int(orc.tables(3)[2])
    terms = 0
    for c in range((lmax + 1) ** 2):
        a, b = product_tables(path, 4, c), orc.tables(4, c)
        assert np.array_equal(a, b), c
        terms += a.size > 0
    assert terms == {"plummer_lmax4.coef": 1, "dehnen_triax_lmax2.coef": 3, "general_lmax2": 9}[case]
    assert lib.agx_debug_tables(path.encode(), 2, 0, None) == 0     # no 2-d tables


@pytest.mark.parametrize("name", FIXTURES + ["ferrers_bar_lmax16.coef"])
def test_cell_polynomials_interpolate_the_node_records(name):
    """The per-cell bi-quintic form used by the FP64 kernels is the tensor-product quintic
    Hermite interpolant of QuinticSpline2d (src/math_spline.cpp:2535-2597): at each of a cell's
    four corners its derivatives d^a/du^a d^b/dv^b (a, b <= 2) equal the corner's node record.
    These 36 conditions determine a bi-quintic uniquely, so they pin the conversion.  Bar: the
    rounding bound of the check itself, 64 eps (sum of the absolute terms + |record|); a wrong
    coefficient misses it by orders of magnitude."""
    path = os.path.join(GOLDEN, name)
    gR, gT = product_tables(path, 0), product_tables(path, 1)
    nr, nt = gR.size, gT.size
    hx, hy = np.diff(gR), np.diff(gT)
    mmax = int(product_tables(path, 3)[3])
    q = {(0, 0): 0, (1, 0): 1, (0, 1): 2, (2, 0): 3, (1, 1): 4, (0, 2): 5, (2, 1): 6, (1, 2): 7, (2, 2): 8}
    eps = np.finfo(float).eps
    tested = 0
    for m in range(-mmax, mmax + 1):
        nodes = product_tables(path, 2, m)
        if nodes.size == 0:
            continue
        rec = nodes.reshape(9, nr, nt)
        c = product_tables(path, 5, m).reshape(nr - 1, nt - 1, 6, 6)
        for ex in (0, 1):
            for ey in (0, 1):
                U = (hx * ex)[:, None]
                V = (hy * ey)[None, :]
                for (da, db), qi in q.items():
                    # d^da/du^da d^db/dv^db of sum c[a][b] u^a v^b at (U, V)
                    val = np.zeros((nr - 1, nt - 1))
                    mag = np.zeros((nr - 1, nt - 1))
                    for a in range(da, 6):
                        for b in range(db, 6):
                            k = np.prod(np.arange(a - da + 1, a + 1)) * np.prod(np.arange(b - db + 1, b + 1))
                            term = k * c[:, :, a, b] * U ** (a - da) * V ** (b - db)
                            val += term
                            mag += np.abs(term)
                    ref = rec[qi, ex:nr - 1 + ex, ey:nt - 1 + ey]
                    bound = 64 * eps * (mag + np.abs(ref))
                    assert np.all(np.abs(val - ref) <= bound), (m, ex, ey, da, db, np.max(np.abs(val - ref) / bound), np.unravel_index(np.argmax(np.abs(val - ref) / bound), val.shape))
                    tested += 1
    assert tested >= 36
