"""Pins of the oracle's DOP853 (oracle/agama_oracle.cpp, restating src/math_ode.cpp:35-423):
bitwise agreement with the reference's own stepper compiled from its source (oracle/_ref),
and the closed-form convergence tests of tests/test_orbit_variational.cpp:36-162."""
import ctypes
import math

import numpy as np
import pytest

from oracle import pyoracle

DP = ctypes.POINTER(ctypes.c_double)


def kepler(t, x, d, ctx):
    r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2]
    r3 = r2 * math.sqrt(r2)
    d[0], d[1], d[2] = x[3], x[4], x[5]
    d[3], d[4], d[5] = -x[0] / r3, -x[1] / r3, -x[2] / r3


def linear1(t, x, d, ctx):
    # x'' = -3.25 x + x'  (tests/test_orbit_variational.cpp:47-55)
    d[0] = x[1]
    d[1] = -3.25 * x[0] + 1.0 * x[1]


def true1(t):
    return (math.cos(math.sqrt(3) * t) - 0.5 / math.sqrt(3) * math.sin(math.sqrt(3) * t)) * math.exp(0.5 * t)


def linear2(t, x, d, ctx):
    # time-dependent 2-d system (tests/test_orbit_variational.cpp:57-73)
    t2, t3 = t * t, t * t * t
    t4 = t3 * t
    den = 1 / (t4 + 4 * t + 4)
    a = [den * (-0.25 * t4 - 2 * t2 - 8 * t - 4), den * (2.75 * t2 + 1.5 * t - 2),
         den * (1.5 * t3 + 8 * t + 8), den * (-t4 - 2 * t2 - 6.5 * t - 1)]
    b = [den * (2 * t3 + 1.5 * t2 + 4), den * (-0.75 * t3 + 2 * t + 4),
         den * (-4 * t2 + 3 * t), den * (2 * t3 - 1.5 * t2)]
    d[0], d[1] = x[2], x[3]
    for i in range(2):
        d[2 + i] = sum(a[i * 2 + k] * x[k] + b[i * 2 + k] * x[k + 2] for k in range(2))


def true2(t):
    return np.array([math.cos(t) + t * math.sin(t / 2), t * math.sin(t) + 2 * math.sin(t / 2)])


def run_adaptive(lib, prefix, fn, x0, T, acc, tsamp):
    cb = pyoracle.RHS(fn)
    n = len(x0)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    out = np.empty(n)
    smp = np.empty(len(tsamp) * n)
    ns = ctypes.c_long()
    rc = getattr(lib, prefix + "_ode_dop853")(n, ctypes.cast(cb, ctypes.c_void_p), None, x0.ctypes.data, T, acc,
                                                out.ctypes.data, len(tsamp), tsamp.ctypes.data, smp.ctypes.data,
                                                ctypes.byref(ns))
    assert rc == 0
    return out, smp.reshape(len(tsamp), n), ns.value


def run_fixed(lib, prefix, fn, x0, h, nsteps):
    cb = pyoracle.RHS(fn)
    n = len(x0)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    st = np.empty(nsteps * n)
    mid = np.empty(nsteps * n)
    rc = getattr(lib, prefix + "_ode_dop853_fixed")(n, ctypes.cast(cb, ctypes.c_void_p), None, x0.ctypes.data,
                                                      h, nsteps, st.ctypes.data, mid.ctypes.data)
    assert rc == 0
    return st.reshape(nsteps, n), mid.reshape(nsteps, n)


@pytest.fixture(scope="module")
def ref():
    r = pyoracle.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return r


@pytest.mark.parametrize("acc", [1e-6, 1e-8, 1e-10])
def test_bitwise_vs_reference_stepper_kepler(ref, acc):
    x0 = [1.0, 0.0, 0.0, 0.0, 1.2, 0.1]     # eccentric Kepler orbit
    ts = np.linspace(0.3, 40.0, 37)
    a = run_adaptive(pyoracle.lib, "orc", kepler, x0, 40.0, acc, ts)
    b = run_adaptive(ref, "ref", kepler, x0, 40.0, acc, ts)
    assert a[2] == b[2]                      # identical step sequence
    assert np.array_equal(a[0], b[0])        # identical end state
    assert np.array_equal(a[1], b[1])        # identical dense output


def test_bitwise_vs_reference_stepper_fixed_linear(ref):
    for fn, x0 in ((linear1, [1.0, 0.0]), (linear2, [1.0, 0.0, 0.0, 1.0])):
        a = run_fixed(pyoracle.lib, "orc", fn, x0, 0.125, 24)
        b = run_fixed(ref, "ref", fn, x0, 0.125, 24)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("system", [1, 2])
def test_fixed_step_convergence_orders(system):
    """end-of-step error ~ dt^8, dense output ~ dt^7 with the reference's bounds"""
    fn, x0, truth = (linear1, [1.0, 0.0], lambda t: np.array([true1(t)])) if system == 1 else \
        (linear2, [1.0, 0.0, 0.0, 1.0], true2)
    nd = 1 if system == 1 else 2
    cs, ci = (2e-4, 3e-4) if system == 1 else (2e-5, 6e-6)
    for k in range(5):
        dt = 0.5 ** k
        nsteps = int(round(3.0 / dt))
        st, mid = run_fixed(pyoracle.lib, "orc", fn, x0, dt, nsteps)
        errs = erri = 0.0
        for s in range(nsteps):
            t = (s + 1) * dt
            if t >= 2.0:
                errs = max(errs, np.max(np.abs(st[s, :nd] - truth(t))))
                erri = max(erri, np.max(np.abs(mid[s, :nd] - truth(t - 0.5 * dt))))
        assert errs < cs * dt ** 8 + 1e-13, (dt, errs)
        assert erri < ci * dt ** 7 + 1e-13, (dt, erri)


def test_adaptive_kepler_energy_and_accuracy():
    """accuracy 1e-10 over ~6 periods: energy conserved to ~1e-9, closed-form Kepler period"""
    x0 = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0]      # circular orbit, period 2 pi
    T = 6 * 2 * math.pi
    out, _, ns = run_adaptive(pyoracle.lib, "orc", kepler, x0, T, 1e-10, np.array([]))
    assert np.max(np.abs(out[:3] - np.array([1.0, 0.0, 0.0]))) < 1e-8
    E0 = 0.5 - 1.0
    E = 0.5 * np.sum(out[3:] ** 2) - 1 / np.linalg.norm(out[:3])
    assert abs(E / E0 - 1) < 1e-9
