"""GPU parity of target plug-ins (orbit(targets=[Target(type='DensitySphHarm', ...)])) against
the oracle, on the triaxial Dehnen Multipole (lmax=8).  Bars: per orbit, the largest
difference of the float32 datacube relative to its largest value is at most 1e-6 for the
median orbit and 1e-5 for 90% of them (the two integrations differ by rounding, chaotic orbits
exponentially; the float32 output itself rounds at 6e-8); the l=0 values of every orbit that
stays inside the grid sum to 1 (partition of unity); zero-length orbits record nothing; two
targets in one call equal one call each."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
GRIDR = np.concatenate([[0.0], np.logspace(-2, 3, 16)])


@pytest.fixture(scope="module")
def setup():
    import agama_amd
    from oracle.pyoracle import OraclePotential
    return agama_amd.Potential(file=COEF), OraclePotential(COEF)


@pytest.mark.parametrize("Omega", [0.0, 0.3])
def test_sphharm_target_matches_oracle(setup, Omega):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(64, lambda p: orc.eval(p)[0], gpu.coefs(), seed=61, periods=5)
    print(f"case: DensitySphHarm target, 64 orbits x 5 periods, Omega={Omega}", flush=True)
    tgt = agama_amd.Target(type="DensitySphHarm", gridr=GRIDR, lmax=6, mmax=4)
    out = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=tgt, Omega=Omega)
    ref = orc.orbits_target_sphharm(ic, T, GRIDR, 6, 4, Omega=Omega)
    assert out.shape == ref.shape == (64, len(tgt)) and out.dtype == np.float32
    rel = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
    print(f"per-orbit max rel diff: median {np.median(rel):.2e}, 90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) <= 1e-6 and np.quantile(rel, 0.9) <= 1e-5
    l0 = out[:, 0] + out[:, 1:17].sum(axis=1)
    assert np.max(np.abs(l0 - 1)) < 2e-6


def test_target_outputs_and_edge_cases(setup):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(8, lambda p: orc.eval(p)[0], gpu.coefs(), seed=62, periods=3)
    print("case: targets with trajectories, two targets, zero length, backward, single orbit", flush=True)
    t1 = agama_amd.Target(type="DensitySphHarm", gridr=GRIDR, lmax=6, mmax=4)
    t2 = agama_amd.Target(type="DensitySphHarm", gridr=np.logspace(-1, 1, 5), lmax=2, mmax=0)
    a, b = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=[t1, t2])
    assert np.array_equal(a, agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=t1))
    assert np.array_equal(b, agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=t2))
    # targets first, then the trajectory (interface_python.cpp:7991-7994)
    m, times, traj = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=t1, trajsize=5, separateTime=True,
                                     dtype="float64")
    assert np.array_equal(m, a) and times.shape == (8, 5) and traj.shape == (8, 5, 6)
    # ... recorded in the same integration: the samples of a plain trajsize=5 call
    t_plain, x_plain = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=5, separateTime=True, dtype="float64")
    assert np.array_equal(times, t_plain)
    assert np.max(np.abs(traj - x_plain)) <= 1e-12 * np.max(np.abs(x_plain))
    z = agama_amd.orbit(potential=gpu, ic=ic, time=0.0, targets=t1)
    assert np.all(z == 0)
    back = agama_amd.orbit(potential=gpu, ic=ic, time=-T, targets=t1)
    assert np.max(np.abs(back[:, 0] + back[:, 1:17].sum(axis=1) - 1)) < 2e-6
    one = agama_amd.orbit(potential=gpu, ic=ic[0], time=T[0], targets=t1)
    assert one.shape == (len(t1),) and np.array_equal(one, a[0])
    with pytest.raises(agama_amd.AgamaAmdError):
        agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=t1, precision=agama_amd.PRECISION_FP32)


@pytest.mark.parametrize("kind", ["DensitySphHarm", "KinemShell"])
def test_targets_with_dprkn8_match_oracle(setup, kind):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(64, lambda p: orc.eval(p)[0], gpu.coefs(), seed=66, periods=5)
    print(f"case: {kind} target with DPRKN8, 64 orbits x 5 periods, Omega=0.2", flush=True)
    if kind == "DensitySphHarm":
        tgt = agama_amd.Target(type=kind, gridr=GRIDR, lmax=6, mmax=4)
        ref = orc.orbits_target_sphharm(ic, T, GRIDR, 6, 4, Omega=0.2, method="dprkn8")
    else:
        gridr = np.concatenate([[0.0], np.logspace(-2, 3, 14)])
        tgt = agama_amd.Target(type=kind, gridr=gridr, degree=2)
        ref = orc.orbits_target_kinem_shell(ic, T, gridr, 2, Omega=0.2, method="dprkn8")
    out = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=tgt, Omega=0.2, method="dprkn8")
    assert out.shape == ref.shape == (64, len(tgt))
    rel = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
    print(f"per-orbit max rel diff: median {np.median(rel):.2e}, 90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) <= 1e-6 and np.quantile(rel, 0.9) <= 1e-5
    # the two integrators sample the same orbits: their target values agree to the integration error
    dop = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=tgt, Omega=0.2)
    d = np.max(np.abs(out - dop), axis=1) / np.max(np.abs(dop), axis=1)
    assert np.median(d) < 1e-2


@pytest.mark.parametrize("degree", [0, 1])
def test_cylindrical_target_matches_oracle(setup, degree):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(64, lambda p: orc.eval(p)[0], gpu.coefs(), seed=63, periods=5)
    gridR = np.concatenate([[0.0], np.logspace(-2, 2, 12)])
    gridz = np.concatenate([[0.0], np.logspace(-2, 2, 10)])
    kind = "DensityCylindricalTopHat" if degree == 0 else "DensityCylindricalLinear"
    print(f"case: {kind} target, 64 orbits x 5 periods", flush=True)
    tgt = agama_amd.Target(type=kind, gridr=gridR, gridz=gridz, mmax=4)
    out = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=tgt, Omega=0.2)
    ref = orc.orbits_target_cyl(ic, T, gridR, gridz, 4, degree, Omega=0.2)
    assert out.shape == ref.shape == (64, len(tgt))
    empty = np.all(ref == 0, axis=1)           # orbits that never enter the grid
    assert np.all(out[empty] == 0)
    rel = np.max(np.abs(out - ref), axis=1)[~empty] / np.max(np.abs(ref), axis=1)[~empty]
    print(f"per-orbit max rel diff: median {np.median(rel):.2e}, 90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) <= 1e-6 and np.quantile(rel, 0.9) <= 1e-5


@pytest.mark.parametrize("degree", [0, 1, 3])
def test_kinem_shell_target_matches_oracle(setup, degree):
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import bspline_values
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(64, lambda p: orc.eval(p)[0], gpu.coefs(), seed=64, periods=5)
    gridr = np.concatenate([[0.0], np.logspace(-2, 3, 14)])
    print(f"case: KinemShell degree {degree} target, 64 orbits x 5 periods, with a density target", flush=True)
    tgt = agama_amd.Target(type="KinemShell", gridr=gridr, degree=degree)
    dens = agama_amd.Target(type="DensitySphHarm", gridr=GRIDR, lmax=4, mmax=2)
    out, d = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=[tgt, dens], Omega=0.2)
    ref = orc.orbits_target_kinem_shell(ic, T, gridr, degree, Omega=0.2)
    assert out.shape == ref.shape == (64, 2 * (len(gridr) + degree - 1)) == (64, len(tgt))
    rel = np.max(np.abs(out - ref), axis=1) / np.max(np.abs(ref), axis=1)
    print(f"per-orbit max rel diff: median {np.median(rel):.2e}, 90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) <= 1e-6 and np.quantile(rel, 0.9) <= 1e-5
    # the density target computed beside it (samples with velocities) is the same as alone
    assert np.array_equal(d, agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=dens, Omega=0.2))
    # the v_r^2 + v_t^2 values summed over the basis: the time average of v^2, positive
    assert np.all(out.sum(axis=1) > 0)
    with pytest.raises(ValueError):
        agama_amd.Target(type="KinemShell", gridr=gridr, degree=4)
    assert bspline_values(degree, gridr, [0.5]).shape == (1, len(gridr) + degree - 1)


@pytest.mark.parametrize("degree", [0, 1])
def test_classic_target_matches_oracle(setup, degree):
    import agama_amd
    from agama_amd import workloads
    gpu, orc = setup
    ic, T = workloads.triaxial_dehnen_orbits(64, lambda p: orc.eval(p)[0], gpu.coefs(), seed=65, periods=5)
    gridr = np.concatenate([[0.0], np.logspace(-2, 2, 12)])
    kind = "DensityClassicTopHat" if degree == 0 else "DensityClassicLinear"
    print(f"case: {kind} target, 64 orbits x 5 periods, 3 strips per pane, axis ratios 0.8, 0.6", flush=True)
    tgt = agama_amd.Target(type=kind, gridr=gridr, stripsPerPane=3, axisRatioY=0.8, axisRatioZ=0.6)
    out = agama_amd.orbit(potential=gpu, ic=ic, time=T, targets=tgt, Omega=0.2)
    ref = orc.orbits_target_classic(ic, T, gridr, degree, 3, 0.8, 0.6, Omega=0.2)
    assert out.shape == ref.shape == (64, len(tgt))
    empty = np.all(ref == 0, axis=1)
    assert np.all(out[empty] == 0)
    rel = np.max(np.abs(out - ref), axis=1)[~empty] / np.max(np.abs(ref), axis=1)[~empty]
    print(f"per-orbit max rel diff: median {np.median(rel):.2e}, 90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) <= 1e-6 and np.quantile(rel, 0.9) <= 1e-5
