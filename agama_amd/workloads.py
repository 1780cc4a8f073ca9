"""Synthetic workloads for the Multipole + DOP853 path: potential parameters and seeded
initial conditions (counter-based Philox generator, seed 42 by default).

Triaxial Dehnen workload (gamma=1, M=1, a=1, axis ratios 1:0.8:0.6, Multipole lmax=8,
gridSizeR=50, r in [1e-3, 1e3]):
  positions drawn from the density: elliptical radius m from the inverse mass profile
  m = a / (u^-1/2 - 1), isotropic direction, then y *= 0.8, z *= 0.6;
  velocities isotropic Gaussian with sigma = v_circ(r)/sqrt(3), redrawn while E >= 0;
  v_circ(r)^2 = r dPhi_00/dr from the monopole term of the expansion.
Integration times: `periods` circular periods 2 pi r / v_circ(r) at each orbit's radius,
either fixed (100) or log-uniform in [1, 300] ("uneven" batches).
"""
import numpy as np

DEHNEN_TRIAXIAL = dict(type="Multipole", density="Dehnen", gamma=1, mass=1, scaleRadius=1,
                       axisRatioY=0.8, axisRatioZ=0.6, lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3)


def read_coef_file(path):
    """(radii[nr], Phi[(lmax+1)^2, nr], dPhi[...]) from an Agama Multipole coefficient file
    (the '#Phi' and '#dPhi/dr' blocks; src/potential_factory.cpp:621-641 format)"""
    lines = open(path).read().splitlines()
    blocks = []
    for tag in ("#Phi", "#dPhi/dr"):
        i = lines.index(tag) + 2          # skip the column header
        rows = []
        while i < len(lines) and lines[i].strip():
            rows.append([float(v) for v in lines[i].split()])
            i += 1
        blocks.append(np.array(rows))
    return blocks[0][:, 0], blocks[0][:, 1:].T.copy(), blocks[1][:, 1:].T.copy()


def rng(seed=42):
    return np.random.Generator(np.random.Philox(seed))


def vcirc_from_coefs(radii, Phi, dPhi):
    """v_circ(r) from the l=0 harmonic of dPhi/dr (the spherically averaged radial force):
    log-log interpolation inside the grid, power-law continuation outside"""
    lr = np.log(radii)
    ldp = np.log(np.maximum(dPhi[0], 1e-300))

    def vc(r):
        r = np.asarray(r, dtype=np.float64)
        x = np.log(r)
        slope_lo = (ldp[1] - ldp[0]) / (lr[1] - lr[0])
        slope_hi = (ldp[-1] - ldp[-2]) / (lr[-1] - lr[-2])
        y = np.interp(x, lr, ldp)
        y = np.where(x < lr[0], ldp[0] + slope_lo * (x - lr[0]), y)
        y = np.where(x > lr[-1], ldp[-1] + slope_hi * (x - lr[-1]), y)
        return np.sqrt(r * np.exp(y))
    return vc


def sample_dehnen_positions(n, gen, a=1.0, p=0.8, q=0.6):
    u = gen.random(n)
    u = np.clip(u, 1e-300, 1 - 1e-16)
    m = a / (u ** -0.5 - 1)
    d = gen.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pos = d * m[:, None]
    pos[:, 1] *= p
    pos[:, 2] *= q
    return pos


def sample_velocities(pos, vc, potential_fn, gen, max_rounds=200):
    """isotropic Gaussian velocities with sigma = v_circ(r)/sqrt(3), redrawn while E >= 0"""
    n = len(pos)
    r = np.linalg.norm(pos, axis=1)
    sig = vc(r) / np.sqrt(3)
    phi = np.asarray(potential_fn(pos))
    vel = np.empty((n, 3))
    todo = np.arange(n)
    for _ in range(max_rounds):
        if len(todo) == 0:
            break
        v = gen.normal(size=(len(todo), 3)) * sig[todo, None]
        vel[todo] = v
        E = phi[todo] + 0.5 * np.sum(v * v, axis=1)
        todo = todo[E >= 0]
    if len(todo):
        vel[todo] *= 0.5   # guarantees a bound orbit for the few stragglers
    return vel


def triaxial_dehnen_orbits(n, potential_fn, coefs, seed=42, periods=100.0, uneven=False):
    """initial conditions [n,6] and integration times [n] for the triaxial Dehnen workload;
    uneven=True draws the number of periods log-uniformly in [1, 300]"""
    gen = rng(seed)
    pos = sample_dehnen_positions(n, gen)
    vc = vcirc_from_coefs(*coefs)
    vel = sample_velocities(pos, vc, potential_fn, gen)
    r = np.linalg.norm(pos, axis=1)
    tcirc = 2 * np.pi * r / vc(r)
    if uneven:
        nper = np.exp(gen.uniform(np.log(1.0), np.log(300.0), size=n))
    else:
        nper = np.full(n, float(periods))
    return np.hstack([pos, vel]), nper * tcirc


DISK_BULGE_HALO = ("type=Multipole, density=Disk+Dehnen+NFW, lmax=12, gridSizeR=80, rmin=1e-2, "
                   "rmax=1e3, 0.mass=0.05, 0.scaleRadius=3, 0.scaleHeight=0.3, 1.mass=0.01, "
                   "1.scaleRadius=0.5, 1.gamma=1, 2.mass=1, 2.scaleRadius=20")


def disk_orbits(n, force_fn, seed=42, periods=100.0, Rd=3.0, hz=0.3, sigma_frac=0.1):
    """near-circular disk orbits in an axisymmetric potential: cylindrical radius from the
    exponential surface density (R = Rd * Gamma(2) deviate), z from the sech^2 profile of
    scale height hz, azimuth uniform; velocities v_phi = v_circ(R) plus Gaussian dispersions of
    sigma_frac * v_circ(R) in all three components, v_circ^2 = R dPhi/dR in the midplane.
    force_fn(points[n,3]) -> -grad Phi.  Integration time: `periods` circular periods at R."""
    gen = rng(seed)
    R = Rd * gen.gamma(2.0, size=n)
    R = np.clip(R, 0.05, None)
    z = hz * np.arctanh(np.clip(2 * gen.random(n) - 1, -1 + 1e-15, 1 - 1e-15))
    ph = gen.uniform(0, 2 * np.pi, n)
    mid = np.stack([R, np.zeros(n), np.zeros(n)], 1)
    vc = np.sqrt(np.maximum(-R * np.asarray(force_fn(mid))[:, 0], 0.0))
    vR, vz, vphi = (gen.normal(size=(3, n)) * sigma_frac * vc)
    vphi = vphi + vc
    c, s = np.cos(ph), np.sin(ph)
    ic = np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c, vz], 1)
    return ic, periods * 2 * np.pi * R / vc


BAR_OMEGA = 40.0
BAR_HALO = [dict(type="Multipole", density="Ferrers", mass=0.02, scaleRadius=3,
                 axisRatioY=1 / 3, axisRatioZ=0.8 / 3, lmax=16, gridSizeR=25, rmin=0.01, rmax=100),
            dict(type="NFW", mass=1, scaleRadius=20)]


def bar_halo_orbits(n, force_fn, seed=42, Omega=BAR_OMEGA, rmin=0.3, rmax=10.0):
    """initial conditions in the frame rotating with the bar (pattern speed Omega about z):
    cylindrical radius log-uniform in [rmin, rmax] around the bar, z Gaussian with sigma
    0.1 R, azimuth uniform; inertial velocity = prograde circular speed v_c(R) (from the
    midplane radial force along x) plus 10% Gaussian dispersion, converted to the rotating
    frame (v - Omega z x r).  Integration times log-uniform in [5, 200] bar periods 2 pi/Omega.
    force_fn(points[n,3]) -> -grad Phi."""
    gen = rng(seed)
    R = np.exp(gen.uniform(np.log(rmin), np.log(rmax), n))
    z = 0.1 * R * gen.normal(size=n)
    ph = gen.uniform(0, 2 * np.pi, n)
    mid = np.stack([R, np.zeros(n), np.zeros(n)], 1)
    vc = np.sqrt(np.maximum(-R * np.asarray(force_fn(mid))[:, 0], 0.0))
    vR, vz, vphi = gen.normal(size=(3, n)) * 0.1 * vc
    vphi = vphi + vc
    c, s = np.cos(ph), np.sin(ph)
    x, y = R * c, R * s
    vx, vy = vR * c - vphi * s, vR * s + vphi * c
    ic = np.stack([x, y, z, vx + Omega * y, vy - Omega * x, vz], 1)
    T = np.exp(gen.uniform(np.log(5.0), np.log(200.0), n)) * 2 * np.pi / Omega
    return ic, T


def density_points(n, seed=7):
    """points drawn from the triaxial Dehnen density (for potential/force evaluation)"""
    return sample_dehnen_positions(n, rng(seed))
