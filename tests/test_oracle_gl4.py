"""Pins of the oracle's Gauss-Legendre variational stepper Ode2StepperGL4<3>
(oracle/agama_oracle.cpp, restating src/math_ode.cpp:1017-1250): bitwise agreement with the
reference's own stepper compiled from its source (oracle/_ref) on a time-dependent linear
system with both A and B terms and varying step lengths (end states and dense output), and the
order of the method on x'' = -w^2 x (closed form)."""
import ctypes
import math

import numpy as np
import pytest

from oracle import pyoracle


def mat_rotating(t, a, b, ctx):
    """a rotating anisotropic oscillator with a Coriolis-like B term"""
    c, s = math.cos(0.7 * t), math.sin(0.7 * t)
    k1, k2, k3 = 1.3, 0.4, 2.1
    a[0] = -(k1 * c * c + k2 * s * s); a[1] = -(k1 - k2) * c * s; a[2] = 0.05 * s
    a[3] = a[1]; a[4] = -(k1 * s * s + k2 * c * c); a[5] = 0.02
    a[6] = a[2]; a[7] = a[5]; a[8] = -k3
    b[0] = 0; b[1] = 0.6; b[2] = 0
    b[3] = -0.6; b[4] = 0; b[5] = 0.01 * c
    b[6] = 0; b[7] = 0; b[8] = -0.01


def run(lib, name, fn, x0, steps, frac):
    cb = pyoracle.MAT(fn)
    nvec = len(x0) // 6
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    steps = np.ascontiguousarray(steps, dtype=np.float64)
    frac = np.ascontiguousarray(frac, dtype=np.float64)
    st = np.empty(len(steps) * nvec * 6)
    de = np.empty(max(len(steps) * len(frac), 1) * nvec * 6)
    rc = getattr(lib, name)(nvec, ctypes.cast(cb, ctypes.c_void_p), None, x0.ctypes.data, len(steps),
                            steps.ctypes.data, st.ctypes.data, len(frac), frac.ctypes.data, de.ctypes.data)
    assert rc == 0
    return st.reshape(len(steps), nvec * 6), de[:len(steps) * len(frac) * nvec * 6]


@pytest.mark.parametrize("nvec", [1, 6])
def test_gl4_bitwise_equal_to_reference(nvec):
    ref = pyoracle.ref_lib()
    if ref is None or not hasattr(ref, "ref_ode_gl4"):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    x0 = rng.normal(size=nvec * 6)
    steps = rng.uniform(0.01, 0.3, 60)
    frac = np.array([0.0, 0.13, 0.5, 0.77, 1.0])
    a = run(pyoracle.lib, "orc_ode_gl4", mat_rotating, x0, steps, frac)
    b = run(ref, "ref_ode_gl4", mat_rotating, x0, steps, frac)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    # backward steps as well
    a = run(pyoracle.lib, "orc_ode_gl4", mat_rotating, x0, -steps[:20], frac)
    b = run(ref, "ref_ode_gl4", mat_rotating, x0, -steps[:20], frac)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_gl4_order_on_harmonic_oscillator():
    """x'' = -w^2 x, x(0) = 1, x'(0) = 0 in all three dimensions: the end error of an
    order-8 method falls by 2^8 when the step halves; dense output at the midpoint is accurate
    to the interpolation order"""
    w = 1.7

    def mat(t, a, b, ctx):
        for i in range(9):
            a[i] = -w * w if i in (0, 4, 8) else 0.0
            b[i] = 0.0
    x0 = [1.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    T = 6.0
    errs = []
    for n in (20, 40, 80):
        st, de = run(pyoracle.lib, "orc_ode_gl4", mat, x0, np.full(n, T / n), [0.5])
        errs.append(abs(st[-1, 0] - math.cos(w * T)) + abs(st[-1, 3] + w * math.sin(w * T)))
        mid = de.reshape(n, 6)[n // 2, 0]
        tmid = (n // 2 + 0.5) * T / n
        assert abs(mid - math.cos(w * tmid)) < 50 * (T / n) ** 6
    assert errs[0] / errs[1] > 150 and errs[1] / errs[2] > 150
