"""Pins of the oracle's every-step recording (trajsize=0, src/orbit.cpp:38-45) and of its
agama.Orbit restatement (createOrbitObject and Orbit.__call__, src/interface_python.cpp:
6605-6880; QuinticSpline from values and derivatives, src/math_spline.cpp:455-486, 653-718,
1376-1429), against what the mathematics fixes:
  - the quintic spline through values and first derivatives reproduces any cubic exactly (the
    cubic meets every condition that defines it: nodes, continuous third derivative, zero fourth
    derivative at the ends), passes through its nodes and is continued linearly beyond them;
  - the inertial-frame splines of a series recorded in a rotating frame recover a straight
    inertial motion exactly, and the output is rotated back by Omega*t;
  - repeated timestamps are dropped, a backward series is put in time order;
  - every-step rows: the start, then each step's end; as many rows as steps + 1; the last row is
    the trajsize=1 result; between the rows the spline agrees with the integrator's own dense
    output (trajsize=S sampling) to the interpolation error."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN
from oracle.pyoracle import OraclePotential, orbit_interp

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def orc():
    return OraclePotential(COEF)


def _rot(p, ang):
    """inertial -> rotating frame at angle ang (orbit.cpp:119-133)"""
    ca, sa = np.cos(ang), np.sin(ang)
    o = p.copy()
    o[..., 0] = p[..., 0] * ca + p[..., 1] * sa
    o[..., 1] = p[..., 1] * ca - p[..., 0] * sa
    o[..., 3] = p[..., 3] * ca + p[..., 4] * sa
    o[..., 4] = p[..., 4] * ca - p[..., 3] * sa
    return o


def test_spline_reproduces_cubic_and_nodes():
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 10, 40))
    c = rng.normal(size=(3, 4))
    pos = np.stack([np.polyval(c[k], t) for k in range(3)], 1)
    vel = np.stack([np.polyval(np.polyder(c[k]), t) for k in range(3)], 1)
    traj = np.hstack([pos, vel])
    tq = np.concatenate([rng.uniform(t[0], t[-1], 200), t])
    out, spl, nodes, rev = orbit_interp(t, traj, 0.0, tq)
    assert nodes == 40 and not rev
    for k in range(3):
        ex = np.polyval(c[k], tq)
        assert np.max(np.abs(out[:, k] - ex)) <= 1e-11 * np.max(np.abs(ex))
        assert np.max(np.abs(out[:, 3 + k] - np.polyval(np.polyder(c[k]), tq))) <= 1e-10 * np.max(np.abs(vel))
        # second derivative of the cubic
        d2 = np.polyval(np.polyder(c[k], 2), tq)
        assert np.max(np.abs(spl[:, 6 + k] - d2)) <= 1e-8 * np.max(np.abs(d2))
    # node values are returned as given
    on, _, _, _ = orbit_interp(t, traj, 0.0, t)
    assert np.array_equal(on, traj)
    # outside the span: NaN for the Orbit, linear continuation for the splines
    o2, s2, _, _ = orbit_interp(t, traj, 0.0, [t[0] - 1, t[-1] + 2])
    assert np.all(np.isnan(o2))
    assert np.allclose(s2[0, :3], pos[0] - vel[0], rtol=1e-14, atol=1e-14)
    assert np.allclose(s2[1, :3], pos[-1] + 2 * vel[-1], rtol=1e-14, atol=1e-14)
    assert np.array_equal(s2[:, 3:6], np.stack([vel[0], vel[-1]])) and np.all(s2[:, 6:] == 0)


def test_spline_is_not_exact_beyond_cubics():
    """a quartic violates the end condition (zero fourth derivative): the spline differs"""
    t = np.linspace(0, 1, 12)
    traj = np.zeros((12, 6))
    traj[:, 0], traj[:, 3] = t ** 4, 4 * t ** 3
    out, _, _, _ = orbit_interp(t, traj, 0.0, [0.04, 0.5])
    assert abs(out[0, 0] - 0.04 ** 4) > 1e-12


def test_rotating_frame_straight_line():
    """inertial uniform motion recorded in a frame rotating at Omega"""
    Om = 0.7
    t = np.linspace(-3, 4, 30)
    x0, v = np.array([1.0, -2.0, 0.5]), np.array([0.3, 0.8, -0.1])
    inert = np.hstack([x0 + v * t[:, None], np.broadcast_to(v, (30, 3))])
    rows = _rot(inert, Om * t)
    tq = np.linspace(-3, 4, 77)
    out, spl, _, _ = orbit_interp(t, rows, Om, tq)
    ex = np.hstack([x0 + v * tq[:, None], np.broadcast_to(v, (77, 3))])
    assert np.allclose(spl[:, :3], ex[:, :3], rtol=0, atol=1e-13)
    assert np.allclose(out, _rot(ex, Om * tq), rtol=0, atol=1e-13)


def test_backward_and_repeated_timestamps():
    t = np.linspace(0, 5, 21)
    traj = np.random.default_rng(1).normal(size=(21, 6))
    tq = np.linspace(0, 5, 33)
    fwd = orbit_interp(t, traj, 0.0, tq)
    bwd = orbit_interp(t[::-1], traj[::-1], 0.0, tq)
    assert bwd[3] and not fwd[3] and bwd[2] == 21
    assert np.array_equal(fwd[0], bwd[0])
    # a repeated timestamp (zero-length orbit: the start recorded twice) is one node
    o, _, nodes, rev = orbit_interp([2.0, 2.0], np.vstack([traj[0], traj[0]]), 0.0, [2.0, 2.5])
    assert nodes == 1 and not rev
    assert np.array_equal(o[0], traj[0]) and np.all(np.isnan(o[1]))
    # an empty series
    o, s, nodes, _ = orbit_interp(np.empty(0), np.empty((0, 6)), 0.0, [1.0])
    assert nodes == 0 and np.all(np.isnan(o)) and np.all(np.isnan(s))


@pytest.mark.parametrize("method", ["dop853", "dprkn8"])
def test_every_step_rows(orc, method):
    from agama_amd import workloads
    ic, T = workloads.triaxial_dehnen_orbits(8, lambda p: orc.eval(p)[0], workloads.read_coef_file(COEF),
                                             seed=31, periods=3)
    r = orc.orbits_steps(ic, T, method=method)
    one = orc.orbits(ic, T, trajsize=1, method=method)
    assert np.array_equal(r["nsteps"], one["nsteps"])
    assert np.array_equal(r["rows"], one["nsteps"] + 1)
    for i, (t, x) in enumerate(r["series"]):
        assert t[0] == 0 and t[-1] == T[i] and np.all(np.diff(t) > 0)
        if method == "dop853":   # getSol(0) of the first step is its starting point
            assert np.array_equal(x[0], ic[i])
        else:                    # the Taylor series of the first step evaluated back at its start
            assert np.allclose(x[0], ic[i], rtol=1e-9, atol=1e-12)
        assert np.array_equal(x[-1], one["final"][i])
        # between the rows the quintic spline follows the integrator's own dense output
        S = 301
        reg = orc.orbits(ic[i], T[i], trajsize=S, method=method)
        out, _, nodes, _ = orbit_interp(t, x, 0.0, reg["time"])
        assert nodes == len(t)
        # interpolation error of steps taken at accuracy 1e-8: about 1e-7 of the orbit size,
        # a few 1e-6 in the widest steps, up to 1.5e-5 in a short last step where the
        # end condition (no fourth derivative) is furthest from the orbit's
        err = np.max(np.abs(out[:, :3] - reg["traj"][:, :3]), axis=1) / np.max(np.abs(reg["traj"][:, :3]))
        assert np.median(err) <= 1e-6 and np.max(err) <= 5e-5


def test_every_step_zero_length_backward_der(orc):
    from agama_amd import workloads
    ic, T = workloads.triaxial_dehnen_orbits(4, lambda p: orc.eval(p)[0], workloads.read_coef_file(COEF),
                                             seed=32, periods=2)
    z = orc.orbits_steps(ic, 0.0)
    assert np.all(z["rows"] == 2)
    for i, (t, x) in enumerate(z["series"]):
        assert np.all(t == 0) and np.array_equal(x[0], ic[i]) and np.array_equal(x[1], ic[i])
    b = orc.orbits_steps(ic, -T, tstart=1.5)
    for i, (t, x) in enumerate(b["series"]):
        assert t[0] == 1.5 and t[-1] == 1.5 - T[i] and np.all(np.diff(t) < 0)
        _, _, nodes, rev = orbit_interp(t, x, 0.0, [1.0])
        assert rev and nodes == len(t)
    # deviation vectors at every step: the identity first, and the same times as the trajectory
    d = orc.orbits_steps(ic, T, der=True)
    for i, (t, x, dv) in enumerate(d["series"]):
        assert dv.shape == (6, len(t), 6)
        assert np.array_equal(dv[:, 0, :], np.eye(6))
        assert np.array_equal(x, orc.orbits_steps(ic[i], T[i])["series"][0][1])
    # regular-sampling deviation vectors agree with the every-step ones at the last row
    v = orc.orbits_var(ic, T, trajsize=2)
    for i, (t, x, dv) in enumerate(d["series"]):
        assert np.allclose(dv[:, -1, :], v["dev"][i, :, -1, :], rtol=1e-12, atol=1e-12)
