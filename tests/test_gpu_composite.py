"""GPU parity of the two-component composite (Ferrers bar Multipole lmax=16 + analytic NFW
halo) and of orbits in the frame rotating with the bar (pattern speed 40), against the CPU
oracle's composite.  Bars: potential and force 1e-12 relative; orbit positions 1e-7 relative
(median) and a Jacobi-energy drift no worse than the oracle's."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

BAR = os.path.join(GOLDEN, "ferrers_bar_lmax16.coef")


@pytest.fixture(scope="module")
def pots():
    import agama_amd
    from oracle.pyoracle import OraclePotential
    gpu = agama_amd.Potential(dict(type="Multipole", file=BAR), dict(type="NFW", mass=1, scaleRadius=20))
    return gpu, OraclePotential(BAR, nfw=(1.0, 20.0))


def test_composite_eval_matches_oracle(pots):
    gpu, orc = pots
    g = np.random.Generator(np.random.Philox(41))
    n = 200_000
    r = np.exp(g.uniform(np.log(1e-3), np.log(2e3), n))
    d = g.normal(size=(n, 3))
    pts = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    pts[:10] = [[0, 0, 0.5], [0, 0, -2], [1, 0, 0], [0, 1, 0], [3, 0, 0], [0, 0, 0.8],
                [1e-4, 0, 0], [500, 0, 0], [0.3, -0.2, 0.1], [2.9, 0.1, 0.0]]
    phi, grad = gpu._eval(pts, True, True)
    ophi, og = orc.eval(pts)
    ep = np.abs(phi / ophi - 1)
    eg = np.linalg.norm(grad - og, axis=1) / np.linalg.norm(og, axis=1)
    print(f"composite eval: potential err max {ep.max():.2e}, force err max {eg.max():.2e}")
    assert ep.max() <= 1e-12 and eg.max() <= 1e-12


def test_rotating_bar_orbits_match_oracle(pots):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = pots
    ic, T = workloads.bar_halo_orbits(256, lambda p: -orc.eval(p)[1], seed=43)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, Omega=workloads.BAR_OMEGA, info=info)
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8, Omega=workloads.BAR_OMEGA)
    assert np.all(info["status"] == 0) and np.all(ref["status"] == 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / \
        np.linalg.norm(ref["final"][:, :3], axis=1)
    same = info["nsteps"] == ref["nsteps"]
    print(f"bar orbits: position err median {np.median(err):.2e} max {err.max():.2e}; identical step "
          f"counts {same.mean():.3f}; Jacobi drift median {np.median(info['maxde']):.2e} "
          f"(oracle {np.median(ref['maxde']):.2e})")
    assert np.median(err) <= 1e-7
    assert np.median(info["maxde"]) <= 1.5 * np.median(ref["maxde"]) + 1e-13


def test_bar_halo_surrogate_path(pots):
    """the rotating bar + NFW workload on its named path: an MLP surrogate trained on the
    composite (l <= 2 part of the bar plus the NFW halo as its base), integrated in the frame
    rotating at Omega = 40 under the persistent schedule, the FP64 Multipole + NFW judging the
    Jacobi integral.  Every orbit either keeps |E_J/E_J0 - 1| <= 1e-6 with the surrogate force or
    is integrated again with the FP64 kernel and then equals the plain FP64 integration bit for
    bit.  Run with the pilot (8192 orbits: the pilot decides whether the batch stays on the
    surrogate) and without it (1024 orbits, each on the surrogate first)."""
    import agama_amd
    from agama_amd import lib, workloads
    gpu, orc = pots
    print("case: surrogate training on the bar + NFW composite (4M samples, 3000 steps)", flush=True)
    sur = agama_amd.Surrogate.train(gpu, samples=4e6, steps=3000, batch=16384, seed=5)
    ic, T = workloads.bar_halo_orbits(8192, lambda p: -orc.eval(p)[1], seed=44)
    try:
        for pilot, n in ((1, 8192), (0, 1024)):
            print(f"case: surrogate orbits at Omega=40, {n} orbits, pilot {'on' if pilot else 'off'}", flush=True)
            assert lib.agx_set_surrogate_pilot(pilot) == 0
            info_s, info_p = {}, {}
            agama_amd.orbit(potential=sur, ic=ic[:n], time=T[:n], accuracy=1e-8, Omega=workloads.BAR_OMEGA,
                            info=info_s, max_rel_de=1e-6)
            agama_amd.orbit(potential=gpu, ic=ic[:n], time=T[:n], accuracy=1e-8, Omega=workloads.BAR_OMEGA,
                            info=info_p)
            st = info_s["status"]
            fb = (st & 8) != 0
            kept = ~fb
            print(f"pilot {'on' if pilot else 'off'}: {fb.sum()} of {n} orbits fell back to the FP64 kernel, "
                  f"{kept.sum()} kept the surrogate" + (f" (pilot: {info_s['pilot'][1]} of {info_s['pilot'][0]} fell "
                                                        f"back)" if pilot else "") +
                  (f"; their max dE_J/E_J {info_s['maxde'][kept].max():.2e}" if kept.any() else ""))
            assert info_s["n_fallback"] == fb.sum()
            assert np.all((st & ~8) == 0) and np.all(info_p["status"] == 0)
            assert np.array_equal(info_s["final"][fb], info_p["final"][fb])
            assert np.array_equal(info_s["nsteps"][fb], info_p["nsteps"][fb])
            assert np.all(info_s["maxde"][kept] <= 1e-6)
    finally:
        lib.agx_set_surrogate_pilot(1)
