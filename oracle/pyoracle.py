"""ctypes wrapper of the CPU oracle (oracle/build/liboracle*.so) - TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker or the
timed CPU baseline; the agama_amd product never imports it."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_vp = ctypes.c_void_p


def build():
    """compile the oracle (and oracle/_ref when the reference tree is present)"""
    subprocess.check_call(["make", "-s", "-C", HERE])
    subprocess.check_call(["sh", os.path.join(HERE, "ref_build.sh")])


def _load(name):
    path = os.path.join(HERE, "build", name)
    if not os.path.exists(path):
        raise ImportError(f"oracle library missing: {path} (run make -C oracle)")
    lib = ctypes.CDLL(path)
    sig = {
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_multipole_from_file": (_vp, [ctypes.c_char_p]),
        "orc_multipole_from_arrays": (_vp, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
        "orc_free": (None, [_vp]),
        "orc_nfw": (_vp, [ctypes.c_double, ctypes.c_double]),
        "orc_composite": (_vp, [ctypes.c_int, _vp]),
        "orc_multipole_tables": (ctypes.c_long, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
        "orc_eval_many": (None, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp, ctypes.c_int]),
        "orc_eval_hess_many": (None, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp, _vp, ctypes.c_int]),
        "orc_orbit_many": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _vp, _vp, ctypes.c_int]),
        "orc_orbit_many_method": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                                 _vp, _vp, ctypes.c_int, ctypes.c_int]),
        "orc_ode_dprkn8": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                          _vp, ctypes.c_int, _vp, _vp, _vp]),
        "orc_ode_dop853": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                          _vp, ctypes.c_int, _vp, _vp, _vp]),
        "orc_orbit_many_var": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.c_int, ctypes.c_int]),
        "orc_orbit_many_lyap": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_long, _vp, ctypes.c_int, _vp, _vp, _vp,
                                               _vp, ctypes.c_int]),
        "orc_tcirc": (ctypes.c_double, [_vp, ctypes.c_double]),
        "orc_orbit_steps": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_long, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                           _vp, _vp, _vp, _vp, ctypes.c_int]),
        "orc_orbit_interp": (ctypes.c_int, [ctypes.c_long, _vp, _vp, ctypes.c_double, ctypes.c_long, _vp, _vp,
                                            _vp, _vp, _vp]),
        "orc_orbit_target_sphharm": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_long, ctypes.c_int, ctypes.c_int, _vp,
                                                    ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]),
        "orc_orbit_target_cyl": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_long, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                                ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]),
        "orc_orbit_target_kinem_shell": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double,
                                                        ctypes.c_double, ctypes.c_long, ctypes.c_int, _vp,
                                                        ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]),
        "orc_orbit_target_classic": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_long, ctypes.c_int, ctypes.c_int, _vp,
                                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp,
                                                    ctypes.c_int, ctypes.c_int]),
        "orc_bspline_values": (None, [ctypes.c_int, _vp, ctypes.c_int, ctypes.c_long, _vp, _vp]),
        "orc_ode_gl4": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                       _vp, _vp]),
        "orc_ode_dop853_fixed": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                                _vp, _vp]),
        "orc_sincos": (None, [ctypes.c_double, _vp, _vp]),
        "orc_atan2": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_binsearch": (ctypes.c_long, [ctypes.c_double, _vp, ctypes.c_long]),
        "orc_lambertw": (ctypes.c_double, [ctypes.c_double, ctypes.c_int]),
        "orc_gl_nodes": (None, [ctypes.c_int, _vp, _vp]),
        "orc_sphharm_array": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp]),
        "orc_sphharm_array2": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp]),
        "orc_spline2d_eval2": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_trig_multi_angle": (None, [ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp]),
        "orc_spline2d_eval": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                             ctypes.c_long, _vp, _vp, _vp, _vp]),
    }
    for n, (res, args) in sig.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load("liboracle.so")          # strict build: the parity checker
lib_fast = None                      # reference optimisation flags: the timed CPU baseline


def fast():
    global lib_fast
    if lib_fast is None:
        lib_fast = _load("liboracle_fast.so")
    return lib_fast


lib_flops = None                     # strict build with an operation-counting double


def flops():
    """the oracle compiled with every double an operation-counting scalar (flopcount.cpp): the
    same results as the strict build, plus flc_ops(nthreads) / flc_reset(nthreads)"""
    global lib_flops
    if lib_flops is None:
        lib_flops = _load("liboracle_flops.so")
        lib_flops.flc_ops.restype = ctypes.c_longlong
        lib_flops.flc_ops.argtypes = [ctypes.c_int]
        lib_flops.flc_reset.restype = None
        lib_flops.flc_reset.argtypes = [ctypes.c_int]
    return lib_flops


RHS = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
MAT = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def _p(a):
    return None if a is None else a.ctypes.data


class OraclePotential:
    """the oracle's Multipole (from a coefficient file or arrays), optionally composed with an
    analytic NFW halo: nfw=(mass, scaleRadius)"""

    def __init__(self, coef_file=None, coefs=None, library=None, nfw=None):
        self.lib = library or lib
        if coef_file is not None:
            self.h = self.lib.orc_multipole_from_file(str(coef_file).encode())
        else:
            radii, Phi, dPhi = (np.ascontiguousarray(a, dtype=np.float64) for a in coefs)
            lmax = int(round(np.sqrt(Phi.shape[0]))) - 1
            self.h = self.lib.orc_multipole_from_arrays(len(radii), lmax, _p(radii), _p(Phi), _p(dPhi))
        if not self.h:
            raise RuntimeError("oracle: " + self.lib.orc_last_error().decode())
        if nfw is not None:
            halo = self.lib.orc_nfw(float(nfw[0]), float(nfw[1]))
            if not halo:
                self.lib.orc_free(self.h)
                self.h = None
                raise RuntimeError("oracle: " + self.lib.orc_last_error().decode())
            comps = (ctypes.c_void_p * 2)(self.h, halo)
            self.h = self.lib.orc_composite(2, ctypes.cast(comps, ctypes.c_void_p))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_free(self.h)

    def tables(self, which, m=0):
        n = self.lib.orc_multipole_tables(self.h, which, m, None)
        out = np.empty(max(n, 0))
        if n > 0:
            self.lib.orc_multipole_tables(self.h, which, m, _p(out))
        return out

    def eval(self, points, nthreads=0):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        phi = np.empty(len(pts))
        grad = np.empty((len(pts), 3))
        self.lib.orc_eval_many(self.h, len(pts), _p(pts), 0.0, _p(phi), _p(grad), nthreads)
        return phi, grad

    def eval_hess(self, points, nthreads=0):
        """potential, gradient and Hessian {dx2, dy2, dz2, dxdy, dydz, dxdz} of Phi"""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        phi = np.empty(len(pts))
        grad = np.empty((len(pts), 3))
        hess = np.empty((len(pts), 6))
        self.lib.orc_eval_hess_many(self.h, len(pts), _p(pts), 0.0, _p(phi), _p(grad), _p(hess), nthreads)
        return phi, grad, hess

    def orbits(self, ic, ttotal, trajsize=1, tstart=None, accuracy=1e-8, Omega=0.0,
               maxNumSteps=int(1e8), nthreads=0, method="dop853"):
        ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
        n = len(ic)
        tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
        sz = np.ascontiguousarray(np.broadcast_to(np.asarray(trajsize, dtype=np.int32), (n,)))
        ts = None if tstart is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
        offs = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64)
        rows = int(sz.sum())
        out = dict(time=np.empty(rows), traj=np.empty((rows, 6)), final=np.empty((n, 6)),
                   nsteps=np.empty(n, np.int64), nattempts=np.empty(n, np.int64),
                   nevals=np.empty(n, np.int64), maxde=np.empty(n), status=np.empty(n, np.int32))
        m = {"dop853": 0, "dprkn8": 1}[method.lower()]
        rc = self.lib.orc_orbit_many_method(self.h, n, _p(ic), _p(ts), _p(tt), _p(sz), _p(offs), accuracy,
                                            Omega, maxNumSteps, _p(out["time"]), _p(out["traj"]),
                                            _p(out["final"]), _p(out["nsteps"]), _p(out["nattempts"]),
                                            _p(out["nevals"]), _p(out["maxde"]), _p(out["status"]), nthreads, m)
        if rc:
            raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
        out["offsets"] = offs
        return out


def _orbits_var(self, ic, ttotal, trajsize=1, tstart=None, accuracy=1e-8, Omega=0.0,
                maxNumSteps=int(1e8), nthreads=0, method="dop853"):
    """orbits with the six deviation vectors of agama.orbit(der=True): dict with time, traj,
    final, dev (n, 6, S, 6) for equal trajsize, nsteps"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    sz = np.ascontiguousarray(np.broadcast_to(np.asarray(trajsize, dtype=np.int32), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    offs = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64)
    rows = int(sz.sum())
    out = dict(time=np.empty(rows), traj=np.empty((rows, 6)), final=np.empty((n, 6)),
               dev=np.empty(rows * 36), nsteps=np.empty(n, np.int64), nlogdev=np.empty(n, np.int64))
    m = {"dop853": 0, "dprkn8": 1}[method.lower()]
    rc = self.lib.orc_orbit_many_var(self.h, n, _p(ic), _p(ts), _p(tt), _p(sz), _p(offs), accuracy, Omega,
                                     maxNumSteps, _p(out["time"]), _p(out["traj"]), _p(out["final"]),
                                     _p(out["dev"]), _p(out["nsteps"]), _p(out["nlogdev"]), nthreads, m)
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    if np.all(sz == sz[0]):
        out["dev"] = out["dev"].reshape(n, 6, int(sz[0]), 6)
    return out


OraclePotential.orbits_var = _orbits_var

MAX_LYAP_BINS = 128


def _orbits_lyap(self, ic, ttotal, tstart=None, accuracy=1e-8, Omega=0.0, maxNumSteps=int(1e8), der=False,
                 nthreads=0):
    """agama.orbit(lyapunov=True[, der=True]) as the oracle restates it: per orbit the
    (lambda * T_circ, t_chaos / T_circ) pair, with the bins, period and last time behind it"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    sz = np.ones(n, np.int32)
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    offs = np.arange(n, dtype=np.int64)
    dev = np.empty(n * 36) if der else None
    bins = np.empty((n, 2, MAX_LYAP_BINS))
    nb = np.empty(n, np.int32)
    period = np.empty(n)
    last = np.empty(n)
    rc = self.lib.orc_orbit_many_lyap(self.h, n, _p(ic), _p(ts), _p(tt), _p(sz), _p(offs), accuracy, Omega,
                                      maxNumSteps, _p(dev), MAX_LYAP_BINS, _p(bins), _p(nb), _p(period),
                                      _p(last), nthreads)
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    out = np.array([lyapunov_from_bins(bins[i, 0, :nb[i]], bins[i, 1, :nb[i]], last[i], period[i])
                    for i in range(n)])
    return dict(lyap=out, bins=bins, nbins=nb, period=period, last=last)


OraclePotential.orbits_lyap = _orbits_lyap


def _orbits_steps(self, ic, ttotal, tstart=None, accuracy=1e-8, Omega=0.0, maxNumSteps=int(1e8),
                  method="dop853", der=False, nthreads=0, capacity=0):
    """agama.orbit(trajsize=0) as the oracle restates it: every step recorded; returns per orbit
    (times, traj[rows, 6]) or, with der, (times, traj, dev[6, rows, 6]), and the step counts.
    The first pass records up to `capacity` rows per orbit (0: it only counts them); orbits with
    more rows are integrated again with blocks of their size."""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    m = {"dop853": 0, "dprkn8": 1}[method.lower()]

    def run(sel, cap):
        k = len(sel)
        offs = np.concatenate([[0], np.cumsum(cap)[:-1]]).astype(np.int64)
        total = int(cap.sum())
        t = np.empty(total)
        x = np.empty((total, 6))
        d = np.empty(total * 36) if der else None
        rows = np.empty(k, np.int64)
        nsteps = np.empty(k, np.int64)
        status = np.empty(k, np.int32)
        ics = np.ascontiguousarray(ic[sel])
        tts = np.ascontiguousarray(tt[sel])
        tss = None if ts is None else np.ascontiguousarray(ts[sel])
        rc = self.lib.orc_orbit_steps(self.h, k, _p(ics), _p(tss), _p(tts), _p(cap), _p(offs), accuracy, Omega,
                                      maxNumSteps, m, int(der), _p(t), _p(x), _p(d), _p(rows), _p(nsteps),
                                      _p(status), nthreads)
        if rc:
            raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
        items = []
        for i in range(k):
            a, r = offs[i], min(rows[i], cap[i])
            item = (t[a:a + r], x[a:a + r])
            if der:
                item += (d[a * 36:a * 36 + 36 * cap[i]].reshape(6, cap[i], 6)[:, :r],)
            items.append(item)
        return items, rows, nsteps, status

    sel = np.arange(n)
    series, rows, nsteps, status = run(sel, np.full(n, int(capacity), np.int64))
    redo = np.nonzero(rows > capacity)[0]
    if len(redo):
        s2, r2, n2, st2 = run(redo, rows[redo].astype(np.int64))
        for j, i in enumerate(redo):
            series[i] = s2[j]
        rows[redo], nsteps[redo], status[redo] = r2, n2, st2
    return dict(series=series, rows=rows, nsteps=nsteps, status=status)


OraclePotential.orbits_steps = _orbits_steps


def sphharm_target_size(lmax, mmax, gridr):
    ng = len(gridr) - (1 if gridr[0] == 0 else 0)
    return ((lmax // 2 + 1) * (mmax // 2 + 1) - (mmax // 2) * (mmax // 2 + 1) // 2) * ng + 1


# integrator names of agama.orbit(method=...) -> the oracle's method codes
_METHODS = {"dop853": 0, "dprkn8": 1}


def _orbits_target_sphharm(self, ic, ttotal, gridr, lmax, mmax, tstart=None, accuracy=1e-8, Omega=0.0,
                           maxNumSteps=int(1e8), nthreads=0, method="dop853"):
    """agama.orbit(targets=[Target(type='DensitySphHarm', ...)]) as the oracle restates it:
    (n, numCoefs) float32"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    gr = np.ascontiguousarray(gridr, dtype=np.float64)
    out = np.empty((n, sphharm_target_size(lmax, mmax, gr)), dtype=np.float32)
    rc = self.lib.orc_orbit_target_sphharm(self.h, n, _p(ic), _p(ts), _p(tt), accuracy, Omega, maxNumSteps,
                                           int(lmax), int(mmax), _p(gr), len(gr), _p(out), nthreads, _METHODS[method])
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    return out


OraclePotential.orbits_target_sphharm = _orbits_target_sphharm


def cyl_target_size(degree, mmax, gridR, gridz):
    nR = len(gridR) - (1 if gridR[0] == 0 else 0)
    nz = len(gridz) - (1 if gridz[0] == 0 else 0)
    return (nz + degree) * (nR * (mmax // 2 + 1) + degree)


def _orbits_target_cyl(self, ic, ttotal, gridR, gridz, mmax, degree, tstart=None, accuracy=1e-8, Omega=0.0,
                       maxNumSteps=int(1e8), nthreads=0, method="dop853"):
    """orbit(targets=[Target(type='DensityCylindricalTopHat' (degree 0) / 'DensityCylindricalLinear'
    (degree 1), ...)]) as the oracle restates it: (n, numCoefs) float32"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    gR = np.ascontiguousarray(gridR, dtype=np.float64)
    gz = np.ascontiguousarray(gridz, dtype=np.float64)
    out = np.empty((n, cyl_target_size(degree, mmax, gR, gz)), dtype=np.float32)
    rc = self.lib.orc_orbit_target_cyl(self.h, n, _p(ic), _p(ts), _p(tt), accuracy, Omega, maxNumSteps, int(degree),
                                       int(mmax), _p(gR), len(gR), _p(gz), len(gz), _p(out), nthreads, _METHODS[method])
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    return out


OraclePotential.orbits_target_cyl = _orbits_target_cyl


def _orbits_target_kinem_shell(self, ic, ttotal, gridr, degree, tstart=None, accuracy=1e-8, Omega=0.0,
                               maxNumSteps=int(1e8), nthreads=0, method="dop853"):
    """orbit(targets=[Target(type='KinemShell', gridr, degree)]) as the oracle restates it:
    (n, 2 * (len(gridr) + degree - 1)) float32, v_r^2 coefficients then v_t^2"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    gr = np.ascontiguousarray(gridr, dtype=np.float64)
    out = np.empty((n, 2 * (len(gr) + degree - 1)), dtype=np.float32)
    rc = self.lib.orc_orbit_target_kinem_shell(self.h, n, _p(ic), _p(ts), _p(tt), accuracy, Omega, maxNumSteps,
                                               int(degree), _p(gr), len(gr), _p(out), nthreads, _METHODS[method])
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    return out


OraclePotential.orbits_target_kinem_shell = _orbits_target_kinem_shell


def classic_target_size(degree, strips, gridr):
    nsh = len(gridr) - (1 if gridr[0] == 0 else 0)
    return (3 * strips * (strips + degree) + degree) * nsh + degree


def _orbits_target_classic(self, ic, ttotal, gridr, degree, stripsPerPane=2, axisRatioY=1.0, axisRatioZ=1.0,
                           tstart=None, accuracy=1e-8, Omega=0.0, maxNumSteps=int(1e8), nthreads=0, method="dop853"):
    """orbit(targets=[Target(type='DensityClassicTopHat' (degree 0) / 'DensityClassicLinear' (degree 1),
    ...)]) as the oracle restates it: (n, numCoefs) float32"""
    ic = np.ascontiguousarray(ic, dtype=np.float64).reshape(-1, 6)
    n = len(ic)
    tt = np.ascontiguousarray(np.broadcast_to(np.asarray(ttotal, dtype=np.float64), (n,)))
    ts = None if tstart is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tstart, dtype=np.float64), (n,)))
    gr = np.ascontiguousarray(gridr, dtype=np.float64)
    out = np.empty((n, classic_target_size(degree, stripsPerPane, gr)), dtype=np.float32)
    rc = self.lib.orc_orbit_target_classic(self.h, n, _p(ic), _p(ts), _p(tt), accuracy, Omega, maxNumSteps,
                                           int(degree), int(stripsPerPane), _p(gr), len(gr), float(axisRatioY),
                                           float(axisRatioZ), _p(out), nthreads, _METHODS[method])
    if rc:
        raise RuntimeError("oracle orbit: " + self.lib.orc_last_error().decode())
    return out


OraclePotential.orbits_target_classic = _orbits_target_classic


def bspline_values(degree, gridr, x):
    """the oracle's B-spline basis of degree 0..3 on gridr at points x: (len(x), len(gridr)+degree-1)"""
    gr = np.ascontiguousarray(gridr, dtype=np.float64)
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    out = np.empty((len(x), len(gr) + degree - 1))
    lib.orc_bspline_values(int(degree), _p(gr), len(gr), len(x), _p(x), _p(out))
    return out


def orbit_interp(t, traj, Omega, tq):
    """agama.Orbit built from one recorded series (times t, rows traj) and evaluated at tq:
    (rotating-frame points [m, 6] with NaN outside, inertial splines [m, 9] = value, first and
    second derivative of x, y, z, number of nodes, reversed)"""
    t = np.ascontiguousarray(t, dtype=np.float64)
    traj = np.ascontiguousarray(traj, dtype=np.float64).reshape(-1, 6)
    tq = np.ascontiguousarray(np.atleast_1d(tq), dtype=np.float64)
    out = np.empty((len(tq), 6))
    spl = np.empty((len(tq), 9))
    nodes = np.empty(1, np.int64)
    rev = np.empty(1, np.int32)
    lib.orc_orbit_interp(len(t), _p(t), _p(traj), float(Omega), len(tq), _p(tq), _p(out), _p(spl), _p(nodes),
                         _p(rev))
    return out, spl, int(nodes[0]), bool(rev[0])
OraclePotential.tcirc = lambda self, E: self.lib.orc_tcirc(self.h, float(E))


def _logexpm1overx(x):
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        return np.where(np.abs(x) < 1e-2, x * (0.5 + x * (1. / 24 - (1. / 2880) * x * x)),
                        np.where(x < 33, np.log(np.expm1(np.minimum(x, 33)) / np.where(x == 0, 1, x)),
                                 x - np.log(np.abs(x))))


def _derlogexpm1overx(x):
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        return np.where(np.abs(x) < 1e-2, 0.5 + x * (1. / 12 - (1. / 720) * x * x),
                        -1 / np.where(x == 0, 1, x) - 1 / np.expm1(-x))


def lyapunov_from_bins(t, lnw, max_time, period):
    """calcLyapunovExponent's fit and decisions (src/orbit_variational.cpp:406-448) on the bins:
    gsl_fit_linear's running-mean slope/intercept as the starting point, then a MINPACK
    Levenberg-Marquardt (scipy's 'lm', the algorithm behind GSL's lmsder) of the model
    A + ln t + [t >= tch] (F(lambda t) - F(lambda tch)), F(x) = ln((e^x - 1)/x), iterated to
    the least-squares minimum (the reference stops GSL's iterations at a 1e-6 relative change
    of the parameters)"""
    from scipy.optimize import least_squares
    t = np.asarray(t, dtype=np.float64)
    lnw = np.asarray(lnw, dtype=np.float64)
    if not (max_time == max_time) or len(t) < 5:
        return np.array([np.nan, np.nan])
    mx = my = dx2 = dxdy = 0.0
    for i in range(len(t)):
        mx += (t[i] - mx) / (i + 1.0)
        my += (lnw[i] - my) / (i + 1.0)
    for i in range(len(t)):
        dx, dy = t[i] - mx, lnw[i] - my
        dx2 += (dx * dx - dx2) / (i + 1.0)
        dxdy += (dx * dy - dxdy) / (i + 1.0)
    slope = dxdy / dx2
    p0 = np.array([slope, my - mx * slope, max_time * 0.5])

    def fun(p):
        lam, A, tch = p
        r = A + np.log(t) + np.where(t < tch, 0, _logexpm1overx(lam * t) - _logexpm1overx(lam * tch)) - lnw
        return np.where(np.isfinite(r), r, 1e10)

    def jac(p):
        lam, A, tch = p
        J = np.empty((len(t), 3))
        J[:, 0] = np.where(t < tch, 0, (t - tch) * _derlogexpm1overx(lam * t))
        J[:, 1] = 1
        J[:, 2] = np.where(t < tch, 0, -lam * _derlogexpm1overx(lam * tch))
        return J
    res = least_squares(fun, p0, jac=jac, method="lm", xtol=1e-14, ftol=1e-15, gtol=1e-15, max_nfev=2000)
    lam, _, tch = res.x
    if lam < 0 or tch >= t[-2]:
        return np.array([0.0, np.inf])
    if tch < t[0]:
        tch = 0.0
    if float(_logexpm1overx(lam * t[-1]) - _logexpm1overx(lam * tch)) < np.log(2):
        return np.array([0.0, np.inf])
    return np.array([lam * period, tch / period])


def ref_lib():
    """the reference's own DOP853 (oracle/_ref), or None when it was not built"""
    path = os.path.join(HERE, "_ref", "libref_dop853.so")
    if not os.path.exists(path):
        return None
    r = ctypes.CDLL(path)
    r.ref_ode_dop853.restype = ctypes.c_int
    r.ref_ode_dop853.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp,
                                 ctypes.c_int, _vp, _vp, _vp]
    r.ref_ode_dop853_fixed.restype = ctypes.c_int
    r.ref_ode_dop853_fixed.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _vp, _vp]
    if hasattr(r, "ref_ode_dprkn8"):
        r.ref_ode_dprkn8.restype = ctypes.c_int
        r.ref_ode_dprkn8.argtypes = r.ref_ode_dop853.argtypes
    if hasattr(r, "ref_ode_gl4"):
        r.ref_ode_gl4.restype = ctypes.c_int
        r.ref_ode_gl4.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp]
    return r
