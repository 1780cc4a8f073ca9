"""Pins of the oracle's variational equation (RuntimeVariational of src/orbit_variational.cpp,
restated in oracle/agama_oracle.cpp on top of the pinned Hessian and GL4 stepper): the six
deviation vectors of agama.orbit(der=True) are the Jacobian of the flow map, so they must equal
central finite differences of the (pinned) orbit integrator and form a symplectic matrix; the
first sample is the identity; sampling and NaN padding follow the trajectory's."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN
from oracle.pyoracle import OraclePotential

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def orc():
    return OraclePotential(COEF)


def _ics(orc, n, seed, periods):
    from agama_amd import workloads
    return workloads.triaxial_dehnen_orbits(n, lambda p: orc.eval(p)[0], workloads.read_coef_file(COEF),
                                            seed=seed, periods=periods)


def _jacobian(dev_end):
    """dev_end[vec, 6] -> J[i, j] = d state_i / d ic_j"""
    return np.asarray(dev_end).T


@pytest.mark.parametrize("Omega", [0.0, 0.3])
def test_deviation_vectors_are_the_flow_jacobian(orc, Omega):
    ic, T = _ics(orc, 8, seed=5, periods=4)
    r = orc.orbits_var(ic, T, trajsize=5, accuracy=1e-12, Omega=Omega, tstart=0.7)
    eps = 1e-6
    err = np.zeros(len(ic))
    for j in range(6):
        sc = np.linalg.norm(ic[:, :3] if j < 3 else ic[:, 3:], axis=1)
        icp, icm = ic.copy(), ic.copy()
        icp[:, j] += eps * sc
        icm[:, j] -= eps * sc
        rp = orc.orbits(icp, T, accuracy=1e-14, Omega=Omega, tstart=0.7)
        rm = orc.orbits(icm, T, accuracy=1e-14, Omega=Omega, tstart=0.7)
        fd = (rp["final"] - rm["final"]) / (2 * eps * sc[:, None])
        dv = r["dev"][:, j, -1, :]
        err = np.maximum(err, np.max(np.abs(fd - dv), axis=1) / np.max(np.abs(fd), axis=1))
    growth = np.abs(r["dev"][:, :, -1, :]).max(axis=(1, 2))
    print(f"Omega={Omega}: FD error {err}, growth {growth}")
    # strongly unstable orbits make the finite differences nonlinear (their error falls as eps^2)
    assert np.median(err) < 2e-5
    assert np.sum(err < 1e-4) >= 6


def test_deviation_matrix_is_symplectic(orc):
    """at step ends (the last sample).  Samples inside a step come from the stepper's dense
    output, whose velocity part is only 4th order: Ode2StepperGL4::getSol
    (math_ode.cpp:1054) weights the s coefficient by 4*5 where the derivative of T^5 s gives 5;
    the restatement keeps that as it is."""
    ic, T = _ics(orc, 16, seed=6, periods=10)
    O = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    for Omega in (0.0, 0.3):
        r = orc.orbits_var(ic, T, trajsize=3, accuracy=1e-10, Omega=Omega)
        for k in range(len(ic)):
            J = _jacobian(r["dev"][k, :, -1, :])
            S = J.T @ O @ J
            scale = np.max(np.abs(J)) ** 2
            assert np.max(np.abs(S - O)) < 1e-10 * max(scale, 1), (k, Omega)


def test_sampling_identity_and_padding(orc):
    ic, T = _ics(orc, 6, seed=7, periods=3)
    r5 = orc.orbits_var(ic, T, trajsize=5)
    r1 = orc.orbits_var(ic, T, trajsize=1)
    # the first sample is the identity; trajsize changes the sampling only
    assert np.array_equal(r5["dev"][:, :, 0, :], np.broadcast_to(np.eye(6), (6, 6, 6)))
    assert np.array_equal(r1["dev"][:, :, 0, :], r5["dev"][:, :, -1, :])
    assert np.array_equal(r5["traj"], orc.orbits(ic, T, trajsize=5)["traj"])   # orbit unaffected
    # a step limit leaves NaN rows where the trajectory has them
    rs = orc.orbits_var(ic, T, trajsize=40, maxNumSteps=5)
    traj_nan = np.isnan(rs["traj"].reshape(6, 40, 6)[:, :, 0])
    assert np.array_equal(np.isnan(rs["dev"][:, 0, :, 0]), traj_nan) and traj_nan.any()
    # zero-length orbits: the identity is recorded twice (sampling interval 0), NaN after
    rz = orc.orbits_var(ic, np.zeros(6), trajsize=4)
    assert np.array_equal(rz["dev"][:, :, 1, :], rz["dev"][:, :, 0, :])
    assert np.all(np.isnan(rz["dev"][:, :, 2:, :]))
