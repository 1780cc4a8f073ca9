"""GPU parity of the NEMO plugin entry points (src/interface_nemo.cpp:74-220) against the oracle,
called through the C ABI as gyrfalcON would: iniacceleration returns the batched accel
function; every combination of numtype 'd'/'f', ndim 3/2 and the flag bits (bit 0 adds the
potential, bit 1 the acceleration, else they are assigned) is checked on 2,000 bodies of which
a third are inactive (flags[i] & 1 == 0: left untouched), with the potential rotating at
Omega = 0.7 (positions rotated by Omega*t with the reference's sincos before the evaluation,
accelerations rotated back).  Bars: 'd' within 1e-12 of the oracle relative to |grad Phi| (and
|Phi|); 'f' within one float ulp of the oracle's double result rounded to float.  Then the older
interface: inipotential, potential_double and potential_float on single bodies."""
import ctypes
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")
ACC = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                       ctypes.c_char)


def nemo_lib():
    from agama_amd import _native
    lib = _native.lib
    lib.iniacceleration.restype = None
    lib.iniacceleration.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ACC),
                                    ctypes.POINTER(ctypes.c_bool), ctypes.POINTER(ctypes.c_bool)]
    lib.inipotential.restype = None
    lib.inipotential.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_char_p]
    for name in ("potential_double", "potential_float"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [ctypes.POINTER(ctypes.c_int)] + [ctypes.c_void_p] * 4
    return lib


def expected(orc, pos, ndim, Omega, t):
    """the oracle's potential and acceleration (double) for bodies at pos[n, ndim]"""
    from oracle.pyoracle import lib as olib
    s, c = np.zeros(1), np.zeros(1)
    olib.orc_sincos(Omega * t, s.ctypes.data, c.ctypes.data)
    s, c = s[0], c[0]
    x, y = pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64)
    z = pos[:, 2].astype(np.float64) if ndim == 3 else np.zeros(len(pos))
    phi, g = orc.eval(np.stack([x * c + y * s, y * c - x * s, z], 1))
    acc = np.stack([-(g[:, 0] * c - g[:, 1] * s), -(g[:, 1] * c + g[:, 0] * s), -g[:, 2]], 1)
    return phi, acc[:, :ndim], np.linalg.norm(g, axis=1)


def test_accel_all_modes_match_oracle():
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    lib = nemo_lib()
    orc = OraclePotential(COEF)
    Omega, t = 0.7, 0.9
    pars = np.array([Omega, 1.0])
    accel = ACC()
    need_m, need_v = ctypes.c_bool(True), ctypes.c_bool(True)
    print("case: iniacceleration(Omega=0.7, G=1)", flush=True)
    lib.iniacceleration(pars.ctypes.data, 2, COEF.encode(), ctypes.byref(accel), ctypes.byref(need_m),
                        ctypes.byref(need_v))
    assert bool(accel) and not need_m.value and not need_v.value
    n = 2000
    pos3 = workloads.density_points(n, seed=51)
    flags = np.where(np.arange(n) % 3 == 0, 2, 1 + 2 * (np.arange(n) % 2)).astype(np.int32)  # 2: inactive
    active = (flags & 1) == 1
    for numtype, dt in ((b"d", np.float64), (b"f", np.float32)):
        for ndim in (3, 2):
            pos = np.ascontiguousarray(pos3[:, :ndim], dtype=dt)
            phi, acc, gnorm = expected(orc, pos, ndim, Omega, t)
            for flag in range(4):
                print(f"case: accel numtype {numtype.decode()} ndim {ndim} flag {flag}", flush=True)
                pot = np.full(n, 0.5, dtype=dt)
                a = np.full((n, ndim), 0.25, dtype=dt)
                accel(ndim, t, n, None, pos.ctypes.data, None, flags.ctypes.data, pot.ctypes.data, a.ctypes.data,
                      flag, numtype)
                # inactive bodies keep what they had
                assert np.all(pot[~active] == dt(0.5)) and np.all(a[~active] == dt(0.25))
                want_p = phi.astype(dt) + dt(0.5) if flag & 1 else phi.astype(dt)
                want_a = acc.astype(dt) + dt(0.25) if flag & 2 else acc.astype(dt)
                if dt is np.float64:
                    ep = np.max(np.abs(pot[active] - want_p[active]) / np.abs(phi[active]))
                    ea = np.max(np.abs(a[active] - want_a[active]).max(axis=1) / gnorm[active])
                    assert ep <= 1e-12 and ea <= 1e-12, (ep, ea)
                else:
                    ulp_p = np.spacing(np.abs(want_p[active]).astype(np.float32))
                    ulp_a = np.spacing(np.abs(want_a[active]).astype(np.float32))
                    assert np.all(np.abs(pot[active] - want_p[active]) <= ulp_p)
                    assert np.all(np.abs(a[active] - want_a[active]) <= ulp_a)
    # NULL flags: every body active
    pos = np.ascontiguousarray(pos3[:100])
    pot, a = np.zeros(100), np.zeros((100, 3))
    accel(3, t, 100, None, pos.ctypes.data, None, None, pot.ctypes.data, a.ctypes.data, 0, b"d")
    phi, acc, gnorm = expected(orc, pos, 3, Omega, t)
    assert np.max(np.abs(a - acc).max(axis=1) / gnorm) <= 1e-12

    # the older interface: one body, assigned, the same rotating potential
    for ndim in (3, 2):
        x = np.ascontiguousarray(pos3[7, :ndim])
        acc1, pot1, tt = np.zeros(ndim), np.zeros(1), np.array([t])
        nd = ctypes.c_int(ndim)
        lib.potential_double(ctypes.byref(nd), x.ctypes.data, acc1.ctypes.data, pot1.ctypes.data, tt.ctypes.data)
        phi, acc, gnorm = expected(orc, x[None, :], ndim, Omega, t)
        assert abs(pot1[0] / phi[0] - 1) <= 1e-12 and np.max(np.abs(acc1 - acc[0])) <= 1e-12 * gnorm[0]
        xf = x.astype(np.float32)
        accf, potf, tf = np.zeros(ndim, np.float32), np.zeros(1, np.float32), np.array([t], np.float32)
        lib.potential_float(ctypes.byref(nd), xf.ctypes.data, accf.ctypes.data, potf.ctypes.data, tf.ctypes.data)
        phi, acc, gnorm = expected(orc, xf[None, :], ndim, Omega, float(tf[0]))
        assert abs(potf[0] - np.float32(phi[0])) <= np.spacing(np.float32(abs(phi[0])))
        assert np.all(np.abs(accf - acc[0].astype(np.float32)) <= np.spacing(np.abs(acc[0]).astype(np.float32)))
    print("case: inipotential without pattern speed", flush=True)
    npar = ctypes.c_int(0)
    lib.inipotential(ctypes.byref(npar), None, COEF.encode())
    x = np.ascontiguousarray(pos3[11])
    acc1, pot1, tt = np.zeros(3), np.zeros(1), np.array([5.0])
    nd = ctypes.c_int(3)
    lib.potential_double(ctypes.byref(nd), x.ctypes.data, acc1.ctypes.data, pot1.ctypes.data, tt.ctypes.data)
    phi, g = orc.eval(x[None, :])
    assert abs(pot1[0] / phi[0] - 1) <= 1e-12 and np.max(np.abs(acc1 + g[0])) <= 1e-12 * np.linalg.norm(g[0])
