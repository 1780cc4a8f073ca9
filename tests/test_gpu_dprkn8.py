"""GPU parity of the persistent DPRKN8 orbit kernel (method='dprkn8') against the CPU oracle's
DPRKN8 (oracle/agama_oracle.cpp, bitwise equal to the reference stepper of
src/math_ode.cpp:426-701), on the triaxial Dehnen Multipole (lmax=8) and on the rotating-bar +
NFW composite.  Bars: regular orbits agree to 1e-7 relative in position after 100 circular
periods at accuracy 1e-8; the GPU energy drift is no worse than the oracle's."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def setup():
    import agama_amd
    from oracle.pyoracle import OraclePotential
    return agama_amd.Potential(file=COEF), OraclePotential(COEF)


def ics(setup, n, seed, periods=100.0, uneven=False):
    from agama_amd import workloads
    gpu, orc = setup
    return workloads.triaxial_dehnen_orbits(n, lambda p: orc.eval(p)[0], gpu.coefs(), seed=seed,
                                            periods=periods, uneven=uneven)


def test_dprkn8_100_periods_match_oracle(setup):
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 256, seed=81)
    print("case: dprkn8 256 orbits x 100 periods, accuracy 1e-8", flush=True)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, method="dprkn8", info=info)
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8, method="dprkn8")
    assert np.all(info["status"] == 0) and np.all(ref["status"] == 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    same_steps = info["nsteps"] == ref["nsteps"]
    print(f"position rel err median {np.median(err):.2e}, 90% {np.quantile(err, 0.9):.2e}, "
          f"max {err.max():.2e}; identical step counts {same_steps.mean():.3f}")
    assert np.median(err) <= 1e-7
    assert np.quantile(err[same_steps], 0.9) <= 1e-7
    assert np.array_equal(info["nattempts"] >= info["nsteps"], np.ones(len(ic), bool))
    assert np.median(info["maxde"]) <= 1.5 * np.median(ref["maxde"]) + 1e-13
    assert np.max(info["maxde"]) < 1e-5
    # the two methods are different integrators: DPRKN8 needs fewer evaluations here
    info5 = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, info=info5)
    assert np.sum(info["nevals"]) < np.sum(info5["nevals"])


def test_dprkn8_trajectory_sampling_matches_oracle(setup):
    """the quintic Taylor dense output (math_ode.cpp:656-701) at 100 sampled times"""
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 64, seed=82, periods=5)
    print("case: dprkn8 trajectory 64 x 100 samples", flush=True)
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=100, dtype="float64", separateTime=True,
                           method="dprkn8")
    ref = orc.orbits(ic, T, trajsize=100, method="dprkn8")
    rt = ref["time"].reshape(64, 100)
    rx = ref["traj"].reshape(64, 100, 6)
    # the first sample is the Taylor series of the first step evaluated back at its start
    # (getSol has no exact branch for offset 0), as in the oracle: close to, not equal to, ic
    assert np.allclose(x[:, 0, :], rx[:, 0, :], rtol=1e-12, atol=1e-15)
    assert np.allclose(x[:, 0, :], ic, rtol=1e-8, atol=1e-10)
    assert np.max(np.abs(t - rt)) <= 1e-15 * np.max(np.abs(rt))
    scale = np.linalg.norm(rx[..., :3], axis=2)
    err = np.linalg.norm(x[..., :3] - rx[..., :3], axis=2) / scale
    verr = np.linalg.norm(x[..., 3:] - rx[..., 3:], axis=2) / np.linalg.norm(rx[..., 3:], axis=2)
    print(f"sample position err median {np.median(err):.2e}, velocity {np.median(verr):.2e}")
    assert np.median(err) < 1e-9 and np.median(verr) < 1e-9


def test_dprkn8_edge_cases(setup):
    """zero-length orbits, the step limit and a single orbit behave as with DOP853"""
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 8, seed=83, periods=3)
    print("case: dprkn8 zero time / maxNumSteps=5 / single orbit", flush=True)
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=0.0, trajsize=4, dtype="float64", separateTime=True,
                           method="dprkn8")
    ref = orc.orbits(ic, np.zeros(8), trajsize=4, method="dprkn8")
    assert np.array_equal(np.isnan(x), np.isnan(ref["traj"].reshape(8, 4, 6)))
    assert np.array_equal(x[:, :2], ref["traj"].reshape(8, 4, 6)[:, :2])
    info = {}
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=50, maxNumSteps=5, dtype="float64",
                           separateTime=True, info=info, method="DPRKN8")
    ref = orc.orbits(ic, T, trajsize=50, maxNumSteps=5, method="dprkn8")
    assert np.all(info["status"] == 2) and np.all(info["nsteps"] == 5)
    assert np.array_equal(np.isnan(t), np.isnan(ref["time"].reshape(8, 50)))
    rx = ref["traj"].reshape(8, 50, 6)
    ok = ~np.isnan(rx[..., 0])
    rel = np.abs(x[ok] - rx[ok]) / np.maximum(np.abs(rx[ok]), 1e-3)
    print(f"step-limited trajectories: max rel diff {rel.max():.2e}")
    assert rel.max() < 1e-9
    t1, x1 = agama_amd.orbit(potential=gpu, ic=ic[0], time=T[0], trajsize=3, separateTime=True, method="dprkn8")
    assert t1.shape == (3,) and x1.shape == (3, 6)


def test_dprkn8_rotating_bar_nfw_composite():
    """rotating-frame integration (eval2 of orbit.cpp:217-250) in the bar + NFW composite"""
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    bar = os.path.join(GOLDEN, "ferrers_bar_lmax16.coef")
    gpu = agama_amd.Potential(dict(type="Multipole", file=bar), dict(type="NFW", mass=1, scaleRadius=20))
    orc = OraclePotential(bar, nfw=(1.0, 20.0))
    ic, T = workloads.bar_halo_orbits(128, lambda p: -orc.eval(p)[1], seed=84)
    print("case: dprkn8 bar+NFW composite, Omega=40, 128 orbits", flush=True)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, Omega=workloads.BAR_OMEGA, method="dprkn8", info=info)
    ref = orc.orbits(ic, T, Omega=workloads.BAR_OMEGA, method="dprkn8")
    assert np.all(info["status"] == 0) and np.all(ref["status"] == 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    print(f"composite position rel err median {np.median(err):.2e}; Jacobi drift median "
          f"{np.median(info['maxde']):.2e} (oracle {np.median(ref['maxde']):.2e})")
    assert np.median(err) <= 1e-7
    assert np.median(info["maxde"]) <= 1.5 * np.median(ref["maxde"]) + 1e-13


def test_dprkn8_uneven_batch(setup):
    """a log-uniform 1..300-period batch through the persistent longest-first queue"""
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 512, seed=85, uneven=True)
    T = T / 20
    print("case: dprkn8 uneven 512-orbit batch", flush=True)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, info=info, method="dprkn8")
    ref = orc.orbits(ic, T, method="dprkn8")
    assert np.all(info["status"] == 0)
    assert np.mean(info["nsteps"] == ref["nsteps"]) > 0.9
