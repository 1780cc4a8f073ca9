"""The batching workload at its full size: 1M orbits of the triaxial Dehnen Multipole (lmax=8)
with log-uniform 1-300 circular periods, in one persistent launch with the default schedule
(probe order, group mode), as bench.py times it.  The oracle cannot integrate the whole batch
within a test, so the batch (DOP853 and DPRKN8) is checked through properties that do not depend on its size:
every orbit completes; the energy errors stay at the integrator's level; an orbit's arithmetic
does not depend on the schedule (orbits integrated on one lane give bitwise the same results in
a small batch of their own, orbits integrated in group mode agree to the integration's accuracy);
and a sample of short orbits matches the oracle under the regular-orbit bars of test_gpu_orbit."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
N = 1_000_000


@pytest.fixture(scope="module", params=["dop853", "dprkn8"])
def batch(request):
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    gpu, orc = agama_amd.Potential(file=COEF), OraclePotential(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(N, lambda p: gpu.potential(p), gpu.coefs(), seed=42, uneven=True)
    method = request.param
    print(f"case: {N} orbits, log-uniform 1-300 periods, default schedule, {method}", flush=True)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=1, separateTime=True, method=method, info=info)
    info["method"] = method
    return gpu, orc, ic, T, info


def energy(gpu, xv):
    return np.asarray(gpu.potential(xv[:, :3]), np.float64) + 0.5 * np.sum(xv[:, 3:] ** 2, axis=1)


def test_full_batch_completes_and_conserves_energy(batch):
    gpu, _, ic, T, info = batch
    assert np.all(info["status"] == 0)
    assert np.all(info["nsteps"] > 0)
    assert np.all(np.isfinite(info["final"]))
    de = np.abs(energy(gpu, info["final"]) / energy(gpu, ic) - 1)
    maxde = info["maxde"]
    q = np.quantile(maxde, [0.5, 0.999])
    print(f"accepted steps {info['nsteps'].sum()}; max dE/E median {q[0]:.2e}, 99.9% {q[1]:.2e}, max "
          f"{maxde.max():.2e}; final dE/E median {np.median(de):.2e}, max {de.max():.2e}")
    # accuracy 1e-8 over 1-300 periods (the sample test below compares with the oracle's)
    assert q[0] <= 3e-7 and q[1] <= 1e-4 and maxde.max() <= 1e-3
    assert np.median(de) <= 2 * q[0] + 1e-13


def test_schedule_does_not_change_an_orbit(batch):
    """256 short orbits (at most 2,000 steps) of the full batch, integrated again as a batch of
    their own (below the probe's and group mode's minimum batch: one lane per orbit, static
    order): the one-lane results are the same to the bit; an orbit that the full batch
    integrated in group mode (its probe ranked it among the longest) agrees to the integration's
    accuracy"""
    import agama_amd
    gpu, _, ic, T, info = batch
    rng = np.random.default_rng(7)
    pick = rng.choice(np.flatnonzero(info["nsteps"] <= 2000), 256, replace=False)
    again = {}
    agama_amd.orbit(potential=gpu, ic=ic[pick], time=T[pick], trajsize=1, separateTime=True,
                    method=info["method"], info=again)
    same = np.all(again["final"] == info["final"][pick], axis=1) & (again["nsteps"] == info["nsteps"][pick])
    other = ~same
    rel = np.linalg.norm(again["final"][:, :3] - info["final"][pick, :3], axis=1) / \
        np.linalg.norm(info["final"][pick, :3], axis=1)
    print(f"bitwise equal: {same.sum()} of 256; the others' largest position difference "
          f"{rel[other].max() if other.any() else 0:.1e}")
    assert same.sum() >= 248
    # a rounding-level difference may flip one accept/reject decision: then the two differ by
    # up to the integration's own error (~1e-8 per step)
    assert np.all(rel[other] <= 1e-4)


def test_full_batch_sample_matches_oracle(batch):
    """64 short orbits of the full batch against the oracle (their own integration times):
    DOP853 under the regular-orbit bars (regular or chaotic from the library's lyapunov=True),
    DPRKN8 under the bars of test_gpu_dprkn8"""
    import agama_amd
    from tests.test_gpu_orbit import regular_orbit_parity
    gpu, orc, ic, T, info = batch
    rng = np.random.default_rng(8)
    pick = rng.choice(np.flatnonzero(info["nsteps"] <= 2000), 64, replace=False)
    if info["method"] == "dop853":
        lyap = np.asarray(agama_amd.orbit(potential=gpu, ic=ic[pick], time=T[pick], lyapunov=True))
        ref = regular_orbit_parity(info["final"][pick], ic[pick], T[pick], orc, lyap[:, 0] == 0,
                                   "full-batch sample")
    else:
        ref = orc.orbits(ic[pick], T[pick], trajsize=1, accuracy=1e-8, method="dprkn8")
        err = np.linalg.norm(info["final"][pick, :3] - ref["final"][:, :3], axis=1) / \
            np.linalg.norm(ref["final"][:, :3], axis=1)
        print(f"full-batch sample: position difference median {np.median(err):.1e}, max {err.max():.1e}")
        assert np.all(ref["status"] == 0)
        assert np.median(err) <= 1e-7
        assert np.quantile(err[info["nsteps"][pick] == ref["nsteps"]], 0.9) <= 1e-7
    same = np.mean(info["nsteps"][pick] == ref["nsteps"])
    print(f"identical step counts {same:.3f}; max dE/E median GPU {np.median(info['maxde'][pick]):.2e} oracle "
          f"{np.median(ref['maxde']):.2e}")
    assert same >= 0.85
    assert np.median(info["maxde"][pick]) <= 1.5 * np.median(ref["maxde"]) + 1e-13
