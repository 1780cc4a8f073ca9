"""GPU parity of the variational equation (orbitVarKernel: agama.orbit(der=True), the six
deviation vectors of RuntimeVariational advanced with the GL4 stepper) against the CPU oracle's
restatement, on the triaxial Dehnen Multipole (lmax=8), in the inertial and a rotating frame.
Bars: the deviation vectors of regular orbits agree to 1e-7 relative to their largest
component (median over orbits; chaotic orbits diverge from any rounding difference); the
trajectory itself is unaffected (same bar as the orbit tests); sampling, NaN padding and the
zero-length case follow the oracle exactly."""
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


def ics(setup, n, seed, periods):
    from agama_amd import workloads
    gpu, orc = setup
    return workloads.triaxial_dehnen_orbits(n, lambda p: orc.eval(p)[0], gpu.coefs(), seed=seed, periods=periods)


@pytest.mark.parametrize("Omega", [0.0, 0.3])
def test_deviation_vectors_match_oracle(setup, Omega):
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 128, seed=101, periods=10)
    print(f"case: der=True 128 orbits x 10 periods, Omega={Omega}", flush=True)
    info = {}
    t, x, d = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=5, der=True, Omega=Omega, timestart=0.4,
                              dtype="float64", separateTime=True, info=info)
    ref = orc.orbits_var(ic, T, trajsize=5, Omega=Omega, tstart=0.4)
    assert d.shape == (128, 6, 5, 6) and np.all(info["status"] == 0)
    rd = ref["dev"]
    scale = np.max(np.abs(rd), axis=(1, 2, 3))
    err = np.max(np.abs(d - rd), axis=(1, 2, 3)) / scale
    same = info["nsteps"] == ref["nsteps"]
    print(f"deviation vectors: rel err median {np.median(err):.2e}, 90% {np.quantile(err, 0.9):.2e}; "
          f"identical step counts {same.mean():.3f}")
    assert np.allclose(d[:, :, 0, :], rd[:, :, 0, :], rtol=0, atol=1e-15)   # the identity, rotated and back
    assert np.median(err) <= 1e-7
    assert np.quantile(err[same], 0.75) <= 1e-6
    rx = ref["traj"].reshape(128, 5, 6)
    perr = np.linalg.norm(x[:, -1, :3] - rx[:, -1, :3], axis=1) / np.linalg.norm(rx[:, -1, :3], axis=1)
    assert np.median(perr) <= 1e-7


def test_der_sampling_padding_and_layouts(setup):
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 6, seed=102, periods=3)
    print("case: der=True step limit / zero length / layouts", flush=True)
    # step limit: NaN rows exactly where the oracle has them
    t, x, d = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=40, maxNumSteps=5, der=True,
                              dtype="float64", separateTime=True)
    ref = orc.orbits_var(ic, T, trajsize=40, maxNumSteps=5)
    assert np.array_equal(np.isnan(d), np.isnan(ref["dev"]))
    for k in range(6):   # values: within 1e-8 of each orbit's largest component
        ok = ~np.isnan(ref["dev"][k])
        assert np.max(np.abs(d[k][ok] - ref["dev"][k][ok])) <= 1e-8 * np.max(np.abs(ref["dev"][k][ok]))
    # zero-length orbits: the start (identity) recorded twice, NaN after
    t, x, d = agama_amd.orbit(potential=gpu, ic=ic, time=0.0, trajsize=4, der=True, dtype="float64",
                              separateTime=True)
    rz = orc.orbits_var(ic, np.zeros(6), trajsize=4)
    assert np.array_equal(np.isnan(d), np.isnan(rz["dev"]))
    assert np.array_equal(d[:, :, :2, :], rz["dev"][:, :, :2, :])
    # trajsize=1: the deviation vectors at the end
    fin = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=1, der=True, dtype="float64",
                          separateTime=True)[2]
    full = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=3, der=True, dtype="float64",
                           separateTime=True)[2]
    assert fin.shape == (6, 6, 1, 6) and np.allclose(fin[:, :, 0, :], full[:, :, -1, :], rtol=1e-12, atol=1e-14)
    # (N, 2) trajectories + (N, 6) deviation-vector objects without separateTime; one orbit
    trajs, devs = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=7, der=True)
    assert trajs.shape == (6, 2) and devs.shape == (6, 6) and devs[0, 0].shape == (7, 6)
    t1, x1, d1 = agama_amd.orbit(potential=gpu, ic=ic[0], time=T[0], trajsize=3, der=True, separateTime=True)
    assert d1.shape == (6, 3, 6)
    with pytest.raises(RuntimeError):
        agama_amd.orbit(potential=gpu, ic=ic, time=T, der=True)
