"""GPU parity of every-step recording (agama.orbit(trajsize=0)) and of agama.Orbit objects
(dtype=object) against the oracle, on the triaxial Dehnen Multipole (lmax=8).
Bars: the device records steps + 1 rows per orbit, the number of steps matching the oracle's
for at least 90% of a regular sample over 5 periods; the step controller turns rounding
differences of its error estimate (a sum of nearly cancelling terms) into step lengths that
differ by about 1e-7 relative (1e-6 of the duration for the median orbit; a few orbits drift
further), so each row is compared with the oracle's interpolated orbit at the row's own time:
1e-9 relative for the median row (measured 1e-13), and the interpolation bar of the oracle's pins
(5e-5) at most; the first row is the initial condition
exactly (DOP853), the last row the trajsize=1 result bit for bit.  The splines built and evaluated on the
device equal the oracle's restatement on the same rows to 1e-12 relative (values) and 1e-10
(derivatives) in the inertial frame (Omega=0).  In a rotating frame the rows are first rotated
into the inertial frame, where the two builds round differently (the device contracts a*b - c*d
into a fused multiply-add) by an ulp of the position; the spline's second derivatives come from
differences of positions divided by the cube of the step, so that ulp grows into about 1e-10 of
the velocity and 1e-8 of the acceleration between the nodes: the bars there are 1e-11 (values),
1e-8 (derivatives) and 1e-6 (second derivatives).  Node values are returned as recorded.  A second integration of the orbits that
outgrow their first-pass block reproduces the one-pass rows bit for bit."""
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


@pytest.mark.parametrize("method", ["dop853", "dprkn8"])
def test_every_step_rows_match_oracle(setup, method):
    import agama_amd
    from oracle.pyoracle import orbit_interp
    gpu, orc = setup
    ic, T = ics(setup, 64, seed=41, periods=5)
    print(f"case: trajsize=0, 64 orbits x 5 periods, {method}", flush=True)
    info = {}
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, dtype="float64", separateTime=True,
                           method=method, info=info)
    ref = orc.orbits_steps(ic, T, method=method)
    fin = agama_amd.orbit(potential=gpu, ic=ic, time=T, method=method)
    assert t.dtype == object and t.shape == (64,) and x.shape == (64,)
    rows = np.array([len(a) for a in t])
    assert np.array_equal(rows, info["nsteps"] + 1) and np.all(info["status"] == 0)
    same = rows == ref["rows"]
    print(f"identical step counts {same.mean():.3f}")
    assert same.mean() >= 0.9
    errs, dts = [], []
    for i in range(64):
        rt, rx = ref["series"][i]
        assert x[i].shape == (rows[i], 6)
        assert t[i][0] == 0 and t[i][-1] == T[i] and np.all(np.diff(t[i]) > 0)
        assert np.array_equal(x[i][-1], fin[i])
        if method == "dop853":
            assert np.array_equal(x[i][0], ic[i])
        else:   # the Taylor series of the first step evaluated back at its start
            assert np.allclose(x[i][0], rx[0], rtol=1e-12, atol=1e-14)
        if same[i]:
            dts.append(np.max(np.abs(t[i] - rt)) / T[i])
        o, _, _, _ = orbit_interp(rt, rx, 0.0, t[i])
        errs.append(np.max(np.abs(x[i][:, :3] - o[:, :3]), axis=1) / np.max(np.abs(rx[:, :3])))
    errs = np.concatenate(errs)
    print(f"step times rel diff max {max(dts):.2e}; rows vs oracle orbit: median {np.median(errs):.2e}, "
          f"max {errs.max():.2e}")
    assert np.median(dts) <= 1e-6
    assert np.median(errs) <= 1e-9 and errs.max() <= 5e-5
    # the deprecated layout: (N, 2) object array of [times, points]
    res = agama_amd.orbit(potential=gpu, ic=ic[:5], time=T[:5], trajsize=0, dtype="float64", method=method)
    assert res.shape == (5, 2)
    for i in range(5):
        assert np.array_equal(res[i, 0], t[i]) and np.array_equal(res[i, 1], x[i])


def test_second_pass_reproduces_rows(setup):
    """a 512-row budget: 8-row blocks, a pool of 32 chunks of 8 rows that a few orbits finish in,
    and a second integration for the rest"""
    import agama_amd
    from agama_amd._native import lib
    gpu, orc = setup
    ic, T = ics(setup, 48, seed=43, periods=3)
    sz = np.where(np.arange(48) % 3 == 0, 7, 0)   # mixed: regular sampling and every step
    print("case: trajsize=0 with a 512-row budget (pool + second pass), mixed trajsize", flush=True)
    t1, x1 = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=sz, dtype="float64", separateTime=True)
    assert lib.agx_set_record_rows(16 * 32) == 0
    try:
        t2, x2 = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=sz, dtype="float64", separateTime=True)
    finally:
        lib.agx_set_record_rows(0)
    for i in range(48):
        assert np.array_equal(t1[i], t2[i]) and np.array_equal(x1[i], x2[i]), i
    # the regularly sampled orbits equal a plain trajsize=7 integration
    tr, xr = agama_amd.orbit(potential=gpu, ic=ic[::3], time=T[::3], trajsize=7, dtype="float64",
                             separateTime=True)
    for j, i in enumerate(range(0, 48, 3)):
        assert np.array_equal(t1[i], tr[j]) and np.array_equal(x1[i], xr[j])


@pytest.mark.parametrize("Omega", [0.0, 0.3])
def test_orbit_objects_match_oracle(setup, Omega):
    import agama_amd
    from oracle.pyoracle import orbit_interp
    gpu, orc = setup
    ic, T = ics(setup, 24, seed=44, periods=3)
    print(f"case: dtype=object, 24 orbits x 3 periods, Omega={Omega}", flush=True)
    orbs = agama_amd.orbit(potential=gpu, ic=ic, time=T, dtype=object, Omega=Omega)
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, dtype="float64", separateTime=True,
                           Omega=Omega)
    assert orbs.shape == (24,) and orbs.dtype == object
    # inertial-frame rows shared exactly with the oracle (Omega=0), or rotated (Omega != 0)
    bar_v, bar_d, bar_s = (1e-12, 1e-10, (1e-12, 1e-10, 1e-10)) if Omega == 0 else (1e-11, 1e-8, (1e-11, 1e-8, 1e-6))
    worst_v = worst_d = 0.0
    worst_s = [0.0, 0.0, 0.0]
    for i in range(24):
        o = orbs[i]
        assert isinstance(o, agama_amd.Orbit)
        assert len(o) == len(t[i]) and not o.reversed and o.Omega == Omega
        assert o[0] == t[i][0] and o[len(o) - 1] == t[i][-1] and o[-1] == t[i][-1]
        tq = np.concatenate([np.linspace(t[i][0], t[i][-1], 157), t[i][::7]])
        got = o(tq)
        ref, spl, nodes, rev = orbit_interp(t[i], x[i], Omega, tq)
        assert nodes == len(o) and not rev
        scale = np.max(np.abs(x[i][:, :3]))
        vscale = np.max(np.abs(x[i][:, 3:]))
        worst_v = max(worst_v, np.max(np.abs(got[:, :3] - ref[:, :3])) / scale)
        worst_d = max(worst_d, np.max(np.abs(got[:, 3:] - ref[:, 3:])) / vscale)
        # the recorded points themselves
        assert np.allclose(o(t[i]), x[i], rtol=1e-14, atol=1e-15 * scale)
        # the inertial-frame spline members with their derivatives
        for k, s in enumerate((o.x, o.y, o.z)):
            for der in range(3):
                a = s(tq, der=der)
                b = spl[:, der * 3 + k]
                worst_s[der] = max(worst_s[der], np.max(np.abs(a - b)) / max(np.max(np.abs(spl[:, der * 3:der * 3 + 3])), 1e-300))
        # outside the time span: NaN; a scalar argument gives one point
        assert np.all(np.isnan(o(np.array([t[i][0] - 1, t[i][-1] + 1]))))
        assert o(float(t[i][3])).shape == (6,)
        assert repr(o).startswith(f"Orbit with {len(o)} nodes on t=[")
    print(f"device vs oracle splines: value {worst_v:.2e}, derivative {worst_d:.2e}; spline members "
          f"{worst_s[0]:.2e} {worst_s[1]:.2e} {worst_s[2]:.2e}")
    assert worst_v <= bar_v and worst_d <= bar_d
    assert all(w <= b for w, b in zip(worst_s, bar_s))
    with pytest.raises(TypeError):
        agama_amd.Orbit()
    with pytest.raises(IndexError):
        orbs[0][len(orbs[0])]


def test_orbit_objects_backward_zero_length_der(setup):
    import agama_amd
    from oracle.pyoracle import orbit_interp
    gpu, orc = setup
    ic, T = ics(setup, 6, seed=45, periods=2)
    print("case: dtype=object backward / zero length / der, trajsize=0 der arrays", flush=True)
    back = agama_amd.orbit(potential=gpu, ic=ic, time=-T, timestart=2.0, dtype=object)
    ref = orc.orbits_steps(ic, -T, tstart=2.0)
    for i in range(6):
        o = back[i]
        assert o.reversed and o[0] == 2.0 and o[len(o) - 1] == 2.0 - T[i]
        assert len(o) == ref["rows"][i] or abs(len(o) - ref["rows"][i]) <= 2
    z = agama_amd.orbit(potential=gpu, ic=ic, time=0.0, dtype=object)
    for i in range(6):
        assert len(z[i]) == 1 and np.array_equal(z[i](0.0), ic[i]) and np.all(np.isnan(z[i](0.1)))
    single = agama_amd.orbit(potential=gpu, ic=ic[0], time=T[0], dtype=object)
    assert isinstance(single, agama_amd.Orbit)
    # deviation vectors at every step, as arrays and as Orbit objects
    t, x, d = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, der=True, dtype="float64",
                              separateTime=True)
    rd = orc.orbits_steps(ic, T, der=True)
    assert d.shape == (6, 6)
    for i in range(6):
        for k in range(6):
            assert d[i, k].shape == (len(t[i]), 6)
            assert np.array_equal(d[i, k][0], np.eye(6)[k])
        rt, rx, rdv = rd["series"][i]
        assert t[i][-1] == rt[-1]
        for k in range(6):   # at the common end time
            err = np.max(np.abs(d[i, k][-1] - rdv[k][-1])) / np.max(np.abs(rdv[k][-1]))
            assert err <= 1e-6, (i, k, err)
    orbs, dv = agama_amd.orbit(potential=gpu, ic=ic, time=T, dtype=object, der=True)
    assert dv.shape == (6, 6)
    for i in range(6):
        tq = np.linspace(0, T[i], 41)
        for k in range(6):
            got = dv[i, k](tq)
            ref_k, _, nodes, _ = orbit_interp(t[i], d[i, k], 0.0, tq)
            assert nodes == len(dv[i, k])
            assert np.max(np.abs(got - ref_k)) <= 1e-12 * np.max(np.abs(d[i, k])), (i, k)
    # lyapunov together with every-step output: the same exponents as without it
    t2, x2, ly = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, lyapunov=True, dtype="float64",
                                 separateTime=True)
    ly0 = agama_amd.orbit(potential=gpu, ic=ic, time=T, lyapunov=True)
    assert np.array_equal(ly, ly0, equal_nan=True)


def test_orbit_spline_methods_match_oracle(setup):
    """the reference's Spline methods on Orbit.x/y/z (third derivative, ext, integrate with x^n or
    another spline, roots, extrema) against the oracle's quintic spline through the same rows"""
    import agama_amd
    from oracle.pyoracle import orbit_interp
    gpu, orc = setup
    ic, T = ics(setup, 6, seed=45, periods=3)
    print("case: Orbit spline methods, 6 orbits x 3 periods", flush=True)
    orbs = agama_amd.orbit(potential=gpu, ic=ic, time=T, dtype=object)
    t, x = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, dtype="float64", separateTime=True)
    gx, gw = np.polynomial.legendre.leggauss(6)
    for i in range(6):
        o, ti = orbs[i], t[i]
        a, b = ti[:-1, None], ti[1:, None]
        tg = (0.5 * (b - a) * (gx + 1) + a).ravel()            # 6 Gauss points per segment
        wg = (0.5 * (b - a) * gw).ravel()
        spl = orbit_interp(ti, x[i], 0.0, tg)[1]               # oracle x, y, z and derivatives
        scale = np.max(np.abs(x[i][:, :3]))
        for k, s in enumerate((o.x, o.y, o.z)):
            # host evaluation from the nodes equals the oracle's spline
            for der in range(3):
                ref = spl[:, der * 3 + k]
                assert np.max(np.abs(s._host()(tg, der) - ref)) <= 1e-10 * max(np.max(np.abs(ref)), 1e-300)
            # the third derivative: central differences of the oracle's second derivative inside
            # the segments
            mid = 0.5 * (ti[:-1] + ti[1:])
            dl = 1e-3 * np.diff(ti)
            d2p = orbit_interp(ti, x[i], 0.0, mid + dl)[1][:, 6 + k]
            d2m = orbit_interp(ti, x[i], 0.0, mid - dl)[1][:, 6 + k]
            fd = (d2p - d2m) / (2 * dl)
            d3 = s(mid, der=3)
            assert np.max(np.abs(d3 - fd)) <= 1e-4 * np.max(np.abs(d3))
            # integrals: with x^n, and with another coordinate's spline (exact Gauss rules)
            for n in (0, 2):
                assert s.integrate(ti[0], ti[-1], n) == pytest.approx(np.sum(wg * spl[:, k] * tg ** n), rel=1e-12)
            other = (k + 1) % 3
            assert s.integrate(ti[0], ti[-1], (o.x, o.y, o.z)[other]) == \
                pytest.approx(np.sum(wg * spl[:, k] * spl[:, other]), rel=1e-11, abs=1e-12 * scale ** 2 * (ti[-1] - ti[0]))
            assert s.integrate(ti[-1], ti[0]) == -s.integrate(ti[0], ti[-1])
            # extrema: both ends and points of zero derivative; roots: the spline takes the value
            ex = s.extrema()
            assert ex[0] == ti[0] and ex[-1] == ti[-1]
            vscale = np.max(np.abs(x[i][:, 3:]))
            if len(ex) > 2:
                dv = orbit_interp(ti, x[i], 0.0, ex[1:-1])[1][:, 3 + k]
                assert np.max(np.abs(dv)) <= 1e-8 * vscale
            y0 = float(np.mean(x[i][:, k]))
            r = s.roots(y0)
            dense = np.linspace(ti[0], ti[-1], 20 * len(ti))
            vals = orbit_interp(ti, x[i], 0.0, dense)[1][:, k] - y0
            assert len(r) == np.count_nonzero(np.sign(vals[1:]) != np.sign(vals[:-1]))
            if len(r):
                assert np.max(np.abs(orbit_interp(ti, x[i], 0.0, r)[1][:, k] - y0)) <= 1e-10 * scale
            # ext: a value outside the recorded span
            assert s(ti[-1] + 1.0, ext=0.0) == 0.0 and s(ti[0], ext=0.0) == s(ti[0])
            # conv: a Gaussian of width sigma against a fine Gauss rule on the oracle's spline; a
            # zero width gives the value inside the span and 0 outside; another spline as kernel
            if k == 0:
                sig = 0.02 * (ti[-1] - ti[0])
                tq = np.array([ti[0] + 0.3 * (ti[-1] - ti[0]), ti[-1] + 2 * sig])
                got = s(tq, conv=sig)
                gx2, gw2 = np.polynomial.legendre.leggauss(64)
                for tj, gj in zip(tq, got):
                    a2, b2 = max(ti[0], tj - 12 * sig), min(ti[-1], tj + 12 * sig)
                    e2 = np.linspace(a2, b2, 41)
                    yy = (0.5 * np.diff(e2)[:, None] * (gx2 + 1) + e2[:-1, None]).ravel()
                    ww = (0.5 * np.diff(e2)[:, None] * gw2).ravel()
                    sv = orbit_interp(ti, x[i], 0.0, yy)[1][:, 0]
                    ref = np.sum(ww * sv * np.exp(-0.5 * ((tj - yy) / sig) ** 2)) / (np.sqrt(2 * np.pi) * sig)
                    assert gj == pytest.approx(ref, rel=1e-9, abs=1e-12 * scale)
                z = s(np.array([ti[1], ti[-1] + 1.0]), conv=np.array([0.0, 0.0]))
                assert z[0] == pytest.approx(s(ti[1]), rel=1e-13) and z[1] == 0.0
                assert np.isfinite(s(ti[2], conv=o.y))
                with pytest.raises(ValueError):
                    s(ti[1], der=1, conv=sig)
    with pytest.raises(ValueError):
        orbs[0].x(1.0, der=4)
