"""agama_amd: MI355X-native Multipole potential evaluation and batched DOP853 orbit integration.

Python mirror of the reference's interface for this path (Agama's Python extension,
src/interface_python.cpp): `Potential` (type='Multipole' from a coefficient file or from a
density), its `potential`, `force` and `eval` methods, `Density` descriptions for the
construction, and `orbit(potential, ic, time, timestart, trajsize, accuracy, maxNumSteps,
Omega, dtype, separateTime)`.  All evaluation and integration runs in libagama_amd.so on
the GPU; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _native
from ._native import AgamaAmdError, check, lib, ptr

__all__ = ["Potential", "Density", "orbit", "Surrogate", "Target", "Orbit", "AgamaAmdError", "PRECISION_FP64",
           "PRECISION_FP32", "set_orbit_groups"]

PRECISION_FP64 = 0
PRECISION_FP32 = 1
METHOD_DPRKN8 = 0x200      # flag of the orbit entry points (AGX_METHOD_DPRKN8)

# keyword spellings accepted by the reference (case-insensitive) and used by the library
_KEYS = {k.lower(): k for k in ["type", "file", "density", "lmax", "mmax", "gridSizeR", "rmin", "rmax",
                                "symmetry", "mass", "scaleRadius", "gamma", "axisRatioY", "axisRatioZ",
                                "scaleHeight", "surfaceDensity"]}


def _fmt(v):
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    return str(v)


class Density:
    """Description of a density profile used to construct a Multipole expansion, mirroring
    agama.Density(type=..., **params) and the composite form agama.Density(d1, d2, ...).
    Supported types: Dehnen, Hernquist, Ferrers, Disk (exponential, scaleHeight>0), NFW,
    Plummer."""

    def __init__(self, *components, **params):
        if components:
            if params:
                raise TypeError("either components or parameters, not both")
            self.components = []
            for c in components:
                if isinstance(c, dict):
                    c = Density(**c)
                if not isinstance(c, Density):
                    raise TypeError("components must be Density objects or dicts")
                self.components.extend(c.components)
        else:
            self.components = [{k.lower(): v for k, v in params.items()}]
            if "type" not in self.components[0]:
                raise TypeError("Density requires a type")

    def _params(self):
        if len(self.components) == 1:
            c = dict(self.components[0])
            items = [("density", c.pop("type"))] + list(c.items())
            return items
        items = [("density", "+".join(str(c["type"]) for c in self.components))]
        for i, c in enumerate(self.components):
            items += [(f"{i}.{k}", v) for k, v in c.items() if k != "type"]
        return items


class Potential:
    """Multipole potential on the GPU.

    Potential(type='Multipole', file='coefs.ini')
    Potential(type='Multipole', density='Dehnen', gamma=1, axisRatioY=0.8, axisRatioZ=0.6,
              lmax=8, gridSizeR=50, rmin=1e-3, rmax=1e3)
    Potential(type='Multipole', density=Density(...), lmax=..., ...)
    Potential(coefs=(radii, Phi, dPhi))   # arrays of shape [nr], [(lmax+1)^2, nr] x 2
    Potential(dict(type='Multipole', ...), dict(type='NFW', mass=1, scaleRadius=20))
                                          # Composite: one Multipole plus an analytic NFW halo

    A string density name is expanded from the density profile by the Poisson solver
    (the reference instead expands the analytic potential of a named model).
    """

    def __init__(self, *args, **kwargs):
        self._h = None
        if args and all(isinstance(a, dict) for a in args) and not kwargs:
            text = "; ".join(self._text(dict(a)) for a in args)
            h = lib.agx_potential_create(text.encode())
            if not h:
                raise AgamaAmdError("Potential construction failed: " + _native.last_error())
            self._h = h
            return
        if len(args) == 1 and isinstance(args[0], str) and not kwargs:
            kwargs = {"file": args[0]}
        elif args:
            raise TypeError("Potential accepts only named arguments (or a single file name, or "
                            "dictionaries of the components of a composite)")
        coefs = kwargs.pop("coefs", None)
        if coefs is not None:
            radii, Phi, dPhi = (np.ascontiguousarray(a, dtype=np.float64) for a in coefs)
            nc = Phi.shape[0]
            lmax = int(round(np.sqrt(nc))) - 1
            h = lib.agx_potential_from_coefs(len(radii), lmax, ptr(radii), ptr(Phi), ptr(dPhi))
        else:
            h = lib.agx_potential_create(self._text(kwargs).encode())
        if not h:
            raise AgamaAmdError("Potential construction failed: " + _native.last_error())
        self._h = h

    @staticmethod
    def _text(kwargs):
        items = []
        for k, v in kwargs.items():
            kl = k.lower()
            if kl == "density" and isinstance(v, Density):
                items += v._params()
            elif kl in _KEYS:
                items.append((_KEYS[kl], v))
            else:
                raise TypeError(f"unknown argument '{k}'")
        if not any(k == "type" for k, _ in items):
            items.insert(0, ("type", "Multipole"))
        return ", ".join(f"{k}={_fmt(v)}" for k, v in items)

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            lib.agx_potential_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def coefs(self):
        """(radii[nr], Phi[(lmax+1)^2, nr], dPhi[...]) of the expansion"""
        nr, lmax = np.zeros(1, np.int32), np.zeros(1, np.int32)
        check(lib.agx_potential_coefs(self._h, nr.ctypes.data_as(_native._ip), lmax.ctypes.data_as(_native._ip),
                                      None, None, None), "coefs")
        nr, lmax = int(nr[0]), int(lmax[0])
        r = np.empty(nr)
        P = np.empty(((lmax + 1) ** 2, nr))
        dP = np.empty_like(P)
        check(lib.agx_potential_coefs(self._h, None, None, ptr(r), ptr(P), ptr(dP)), "coefs")
        return r, P, dP

    def export(self, filename):
        """write the coefficients in the reference's text format"""
        check(lib.agx_potential_write(self._h, str(filename).encode()), "export")

    def _eval(self, points, want_pot, want_grad, t=0.0, precision=PRECISION_FP64):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        single = pts.ndim == 1
        pts2 = pts.reshape(-1, 3)
        n = pts2.shape[0]
        phi = np.empty(n) if want_pot else None
        grad = np.empty((n, 3)) if want_grad else None
        check(lib.agx_eval_many(self._h, n, ptr(pts2), float(t), ptr(phi), ptr(grad), int(precision)),
              "eval")
        if single:
            phi = phi[0] if phi is not None else None
            grad = grad[0] if grad is not None else None
        return phi, grad

    def potential(self, points, t=0.0, precision=PRECISION_FP64):
        """Phi at points (N x 3 or 3) - reference: Potential.potential"""
        return self._eval(points, True, False, t, precision)[0]

    def force(self, points, t=0.0, precision=PRECISION_FP64):
        """acceleration -grad(Phi) - reference: Potential.force"""
        return -self._eval(points, False, True, t, precision)[1]

    def _eval_hess(self, points, t=0.0):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        single = pts.shape == (3,)
        pts2 = pts.reshape(-1, 3)
        n = pts2.shape[0]
        phi, grad, hess = np.empty(n), np.empty((n, 3)), np.empty((n, 6))
        check(lib.agx_eval_hess_many(self._h, n, ptr(pts2), float(t), ptr(phi), ptr(grad), ptr(hess),
                                     PRECISION_FP64), "eval_hess")
        if single:
            return phi[0], grad[0], hess[0]
        return phi, grad, hess

    def eval(self, points, pot=False, acc=False, der=False, t=0.0, precision=PRECISION_FP64):
        """reference: Potential.eval(points, pot=, acc=, der=) (interface_python.cpp:3114-3207):
        potential, acceleration -dPhi/dx and force derivatives -d2Phi/dxdy in the order
        xx, yy, zz, xy, yz, zx (FP64 when der=True)"""
        if der:
            phi, grad, hess = self._eval_hess(points, t)
        else:
            phi, grad = self._eval(points, pot, acc, t, precision)
        out = []
        if pot:
            out.append(phi)
        if acc:
            out.append(-grad)
        if der:
            out.append(-hess)
        return out[0] if len(out) == 1 else tuple(out)

    def forceDeriv(self, points, t=0.0):
        """reference: Potential.forceDeriv (deprecated there in favour of eval(acc=True, der=True)):
        (force, force derivatives)"""
        _, grad, hess = self._eval_hess(points, t)
        return -grad, -hess


def orbit(potential=None, ic=None, time=None, timestart=None, trajsize=None, accuracy=1e-8,
          maxNumSteps=None, Omega=0.0, dtype="float32", separateTime=False, method="dop853",
          verbose=False, info=None, precision=PRECISION_FP64, max_rel_de=1e-6, der=False, lyapunov=False,
          targets=None):
    """Integrate orbits with DOP853 or DPRKN8 (method=) on the GPU - reference: agama.orbit
    (src/interface_python.cpp:7689-8281, output conventions :7941-7961).

    Returns the trajectory output as in the reference: with separateTime=True and equal
    trajsize a tuple (times[N,S], traj[N,S,6]); with separateTime=False an (N,2) object array
    of [times, traj] pairs; a single orbit (ic of length 6) drops the leading dimension.
    `info`, if a dict, receives per-orbit diagnostics (nsteps, nattempts, nevals, maxde, status,
    final) that the reference does not return.

    `potential` may also be a `Surrogate`: orbits are then integrated with the MLP surrogate
    force while their energy error (judged by the FP64 Multipole) stays within `max_rel_de`
    and they stay within the training radii, and otherwise integrated again with the FP64
    Multipole (status + 8; info["n_fallback"] counts them).

    der=True also integrates the variational equation (src/orbit_variational.cpp) and returns
    the six deviation vectors sampled like the trajectory, as an extra item: an array
    (N, 6, trajsize, 6) with separateTime=True and equal trajsize, otherwise an (N, 6) object
    array of (trajsize, 6) arrays (interface_python.cpp:8035-8047, 8205-8232).
    lyapunov=True adds an (N, 2) array: the Lyapunov exponent in units of the orbital period
    T_circ(E) and the chaos onset time in the same unit (0 and inf for regular orbits,
    orbit_variational.cpp:360-448) as the last item.

    trajsize=0 records every integration step (the start, then each step's end; orbit.cpp:38-45)
    and returns arrays of each orbit's own length.  dtype=object returns agama.Orbit objects
    (`Orbit`: quintic-spline interpolators of the trajectory, interface_python.cpp:6605-6880)
    instead of arrays, recorded at every step unless trajsize is given (:7864-7869); with der=True
    the deviation vectors become an (N, 6) array of Orbit objects too.

    targets= (a Target or a sequence of up to 4) collects each target's data during the
    integration (RuntimeFncTarget, galaxymodel_target.h:95-155) and puts one (N, numCoefs)
    float32 array per target first in the output (interface_python.cpp:7991-7994, 8153-8162).
    Targets are collected with DOP853 or DPRKN8 in FP64.  Regular trajectory samples
    (trajsize > 0) are recorded in the same integration as the targets; every-step output
    (trajsize=0), Orbit objects (dtype=object), deviation vectors (der=True) and Lyapunov
    exponents requested with targets come from a second integration of the same orbits."""
    surrogate = isinstance(potential, Surrogate)
    if not surrogate and not isinstance(potential, Potential):
        raise TypeError("Argument 'potential' must be a valid Potential object")
    # method names as parsed by the reference (interface_python.cpp:7913-7926; stringsEqual is
    # case-insensitive)
    mname = str(method).lower() if isinstance(method, str) else None
    if mname == "hermite":
        raise AgamaAmdError("method='hermite' is not provided by this version")
    if mname not in ("dop853", "dprkn8"):
        raise ValueError("Unknown ODE integation method (valid values are 'dop853', 'dprkn8' or 'hermite')")
    if surrogate and mname != "dop853":
        raise AgamaAmdError("the surrogate path integrates with method='dop853' only")
    method_flag = METHOD_DPRKN8 if mname == "dprkn8" else 0
    if der not in (False, True, 0, 1):
        raise TypeError("Argument 'der' must be a boolean or an int 0/1")
    der = bool(der)
    obj = dtype is object or (isinstance(dtype, str) and dtype in ("object", "O")) or dtype is np.object_
    if der and trajsize is None and not obj:
        raise RuntimeError("Derivatives are requested, but the trajectory itself is not: provide "
                           "trajsize=... for array output or dtype=object for Orbit object output")
    if der and surrogate:
        raise AgamaAmdError("der=True is integrated with the FP64 Multipole only")
    if lyapunov not in (False, True, 0, 1):
        raise TypeError("Argument 'lyapunov' must be a boolean")
    lyapunov = bool(lyapunov)
    if lyapunov and surrogate:
        raise AgamaAmdError("lyapunov=True is integrated with the FP64 Multipole only")
    icarr = np.ascontiguousarray(ic, dtype=np.float64)
    if not (icarr.shape == (6,) or (icarr.ndim == 2 and icarr.shape[1] == 6)):
        raise TypeError("Argument 'ic' does not contain a valid array of length Nx6")
    single = icarr.ndim == 1
    ic2 = icarr.reshape(-1, 6)
    n = ic2.shape[0]
    if time is None:
        raise ValueError("Argument 'time' must be provided")
    tt = np.asarray(time, dtype=np.float64)
    if tt.ndim == 0:
        tt = np.full(n, float(tt))
    elif tt.shape != (n,):
        raise ValueError("Argument 'time' must either be a scalar or have the same length as the number of points in the initial conditions")
    tt = np.ascontiguousarray(tt)
    ts = None
    if timestart is not None:
        ts = np.asarray(timestart, dtype=np.float64)
        ts = np.full(n, float(ts)) if ts.ndim == 0 else np.ascontiguousarray(ts)
        if ts.shape != (n,):
            raise ValueError("Argument 'timestart' must either be a scalar or have the same length as the number of points in the initial conditions")
    have_traj = trajsize is not None
    if have_traj:
        sz = np.asarray(trajsize)
        sz = np.full(n, int(sz), dtype=np.int32) if sz.ndim == 0 else np.ascontiguousarray(sz, dtype=np.int32)
        if sz.shape != (n,) or np.any(sz < 0):
            raise ValueError("Argument 'trajsize', if provided, must either be a nonnegative integer or an array of integers with the same length as the number of points in the initial conditions")
    elif obj:   # Orbit objects without trajsize: every timestep (interface_python.cpp:7864-7869)
        sz = np.zeros(n, dtype=np.int32)
        have_traj = True
    else:
        sz = np.ones(n, dtype=np.int32)
    if obj:
        out_dtype = object
    elif dtype in ("float32", np.float32, "float", float, "float64", np.float64, "double"):
        out_dtype = np.float32 if dtype in ("float32", np.float32) else np.float64
    elif dtype in ("complex64", "complex128", "complex", np.complex64, np.complex128):
        out_dtype = np.complex64 if dtype in ("complex64", np.complex64) else np.complex128
    else:
        raise AgamaAmdError(f"dtype={dtype} is not provided by this version")
    maxsteps = int(1e8) if maxNumSteps is None else int(maxNumSteps)
    if targets is not None:
        tlist = list(targets) if isinstance(targets, (list, tuple)) else [targets]
        if not all(isinstance(t, Target) for t in tlist) or not 1 <= len(tlist) <= 4:
            raise TypeError("Argument 'targets' must be a Target or a sequence of up to 4 Targets")
        if surrogate or precision != PRECISION_FP64:
            raise AgamaAmdError("targets are collected with the FP64 Multipole only")
        mats = [np.empty((n, len(t)), dtype=np.float32) for t in tlist]
        final_t, nsteps_t, status_t = np.empty((n, 6)), np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int32)
        tptr = (ctypes.c_void_p * len(tlist))(*[t._h for t in tlist])
        optr = (ctypes.c_void_p * len(tlist))(*[m.ctypes.data for m in mats])
        if trajsize is not None and not obj and not der and not lyapunov and np.all(sz > 0):
            # regular trajectory samples recorded in the same integration as the targets
            rows = int(sz.sum())
            times, traj = np.empty(rows), np.empty((rows, 6))
            check(lib.agx_orbit_target_traj_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt), ptr(sz),
                                                  float(accuracy), float(Omega), maxsteps, len(tlist), tptr, optr,
                                                  ptr(times), ptr(traj), ptr(final_t), ptr(nsteps_t),
                                                  ptr(status_t), int(precision) | method_flag), "orbit")
            mats = [m[0] if single else m for m in mats]
            if isinstance(info, dict):
                info.update(nsteps=nsteps_t, status=status_t, final=final_t)
            out = _trajectory_outputs(n, single, sz, times, traj, None, out_dtype, separateTime)
            return tuple(mats) + (out if isinstance(out, tuple) else (out,))
        check(lib.agx_orbit_target_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt), float(accuracy),
                                         float(Omega), maxsteps, len(tlist), tptr, optr, ptr(final_t), ptr(nsteps_t),
                                         ptr(status_t), int(precision) | method_flag), "orbit")
        mats = [m[0] if single else m for m in mats]
        if trajsize is None and not obj and not der and not lyapunov:
            if isinstance(info, dict):
                info.update(nsteps=nsteps_t, status=status_t, final=final_t)
            return mats[0] if len(mats) == 1 else tuple(mats)
        rest = orbit(potential=potential, ic=ic, time=time, timestart=timestart, trajsize=trajsize,
                     accuracy=accuracy, maxNumSteps=maxNumSteps, Omega=Omega, dtype=dtype,
                     separateTime=separateTime, method=method, info=info, precision=precision, der=der,
                     lyapunov=lyapunov)
        return tuple(mats) + (rest if isinstance(rest, tuple) else (rest,))
    if have_traj and (obj or np.any(sz == 0)):
        if surrogate:
            raise AgamaAmdError("trajsize=0 and dtype=object are integrated with the FP64 Multipole only")
        return _recorded_orbits(potential, n, single, ic2, ts, tt, sz, accuracy, Omega, maxsteps, method_flag,
                                precision, der, lyapunov, obj, out_dtype, separateTime, info)
    rows = int(sz.sum())
    times = np.empty(rows)
    traj = np.empty((rows, 6))
    final = np.empty((n, 6))
    nsteps = np.empty(n, dtype=np.int64)
    natt = np.empty(n, dtype=np.int64)
    nev = np.empty(n, dtype=np.int64)
    maxde = np.empty(n)
    status = np.empty(n, dtype=np.int32)
    nfb = np.zeros(1, dtype=np.int64)
    if surrogate:
        check(lib.agx_surrogate_orbit_batch(potential._h, n, ptr(ic2), ptr(ts), ptr(tt), ptr(sz), float(accuracy),
                                            float(Omega), maxsteps, float(max_rel_de), ptr(times), ptr(traj),
                                            ptr(final), ptr(nsteps), ptr(natt), ptr(nev), ptr(maxde), ptr(status),
                                            ptr(nfb)), "surrogate orbit")
    elif der or lyapunov:
        dev = np.empty(int(sz.sum()) * 36) if der else None
        lyap = np.empty((n, 2)) if lyapunov else None
        check(lib.agx_orbit_var_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt), ptr(sz), float(accuracy),
                                      float(Omega), maxsteps, ptr(times), ptr(traj), ptr(final), ptr(dev),
                                      ptr(lyap), ptr(nsteps), ptr(natt), ptr(nev), ptr(maxde), ptr(status),
                                      int(precision) | method_flag), "orbit")
    else:
        check(lib.agx_orbit_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt), ptr(sz), float(accuracy),
                                  float(Omega), maxsteps, ptr(times), ptr(traj), ptr(final), ptr(nsteps),
                                  ptr(natt), ptr(nev), ptr(maxde), ptr(status), int(precision) | method_flag),
              "orbit")
    if isinstance(info, dict):
        info.update(nsteps=nsteps, nattempts=natt, nevals=nev, maxde=maxde, status=status, final=final)
        if surrogate:
            info["n_fallback"] = int(nfb[0])
            pil = np.zeros(2, dtype=np.int64)
            lib.agx_surrogate_last_pilot(ptr(pil[:1]), ptr(pil[1:]))
            info["pilot"] = (int(pil[0]), int(pil[1]))   # pilot orbits, of which fell back
    if not have_traj:
        if lyapunov:
            return lyap[0] if single else lyap
        return final[0] if single else final
    out = _trajectory_outputs(n, single, sz, times, traj, dev if der else None, out_dtype, separateTime)
    if lyapunov:
        out = (out if isinstance(out, tuple) else (out,)) + ((lyap[0] if single else lyap),)
    return out


def _recorded_orbits(potential, n, single, ic2, ts, tt, sz, accuracy, Omega, maxsteps, method_flag, precision,
                     der, lyapunov, obj, out_dtype, separateTime, info):
    """orbit() with every-step recording (some trajsize 0) or Orbit-object output: the batch is
    recorded in device memory (agx_orbit_record_batch), then returned as arrays of each orbit's
    length or as Orbit objects over that record (interface_python.cpp:8165-8232)"""
    from .orbitobj import Orbit, _Record
    rows = np.empty(n, dtype=np.int64)
    final = np.empty((n, 6))
    nsteps = np.empty(n, dtype=np.int64)
    natt = np.empty(n, dtype=np.int64)
    nev = np.empty(n, dtype=np.int64)
    maxde = np.empty(n)
    status = np.empty(n, dtype=np.int32)
    h = lib.agx_orbit_record_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt), ptr(sz), float(accuracy),
                                   float(Omega), maxsteps, int(der), ptr(rows), ptr(final), ptr(nsteps), ptr(natt),
                                   ptr(nev), ptr(maxde), ptr(status), int(precision) | method_flag)
    rec = _Record(h, n, der)
    lyap = None
    if lyapunov:   # the exponent from the one-vector variational equation, as without a record
        lyap = np.empty((n, 2))
        check(lib.agx_orbit_var_batch(potential.handle, n, ptr(ic2), ptr(ts), ptr(tt),
                                      ptr(np.ones(n, dtype=np.int32)), float(accuracy), float(Omega), maxsteps,
                                      None, None, None, None, ptr(lyap), None, None, None, None, None,
                                      int(precision) | method_flag), "orbit")
    if isinstance(info, dict):
        info.update(nsteps=nsteps, nattempts=natt, nevals=nev, maxde=maxde, status=status, final=final)
    if obj:
        orbits = np.empty(n, dtype=object)
        for i in range(n):
            orbits[i] = Orbit._make(rec, i, Omega)
        out = orbits[0] if single else orbits
        if der:
            darr = np.empty((n, 6), dtype=object)
            for i in range(n):
                for k in range(6):
                    darr[i, k] = Orbit._make(rec, n + 6 * i + k, Omega)
            out = (out, darr[0] if single else darr)
    else:
        times, traj, dev = rec.fetch(rows, der)
        out = _trajectory_outputs(n, single, rows, times, traj, dev, out_dtype, separateTime, uniform=False)
    if lyapunov:
        out = (out if isinstance(out, tuple) else (out,)) + ((lyap[0] if single else lyap),)
    return out


def _trajectory_outputs(n, single, sz, times, traj, dev, out_dtype, separateTime, uniform=None):
    """the trajectory (and deviation-vector) outputs in the reference's layouts
    (interface_python.cpp:7995-8047); uniform: one trajsize > 0 for all orbits (the 2d/3d
    layout of separateTime=True)"""
    der = dev is not None
    if uniform is None:
        uniform = bool(np.all(sz == sz[0]))

    def conv(a):
        if out_dtype in (np.complex64, np.complex128):
            return (a[..., :3] + 1j * a[..., 3:]).astype(out_dtype)
        return a.astype(out_dtype)

    offs = np.concatenate([[0], np.cumsum(sz)])
    if separateTime and uniform:
        S = int(sz[0])
        t_out = times.reshape(n, S)
        x_out = conv(traj.reshape(n, S, 6))
        if der:
            d_out = conv(dev.reshape(n, 6, S, 6))
            return (t_out[0], x_out[0], d_out[0]) if single else (t_out, x_out, d_out)
        return (t_out[0], x_out[0]) if single else (t_out, x_out)
    parts = [(times[offs[i]:offs[i + 1]], conv(traj[offs[i]:offs[i + 1]])) for i in range(n)]
    darr = None
    if der:   # (N, 6) object array of (trajsize, 6) arrays
        darr = np.empty((n, 6), dtype=object)
        for i in range(n):
            block = dev[offs[i] * 36:offs[i + 1] * 36].reshape(6, sz[i], 6)
            for k in range(6):
                darr[i, k] = conv(block[k])
    if separateTime:
        tarr = np.empty(n, dtype=object)
        xarr = np.empty(n, dtype=object)
        for i, (a, b) in enumerate(parts):
            tarr[i], xarr[i] = a, b
        if der:
            return (tarr[0], xarr[0], darr[0]) if single else (tarr, xarr, darr)
        return (tarr[0], xarr[0]) if single else (tarr, xarr)
    res = np.empty((n, 2), dtype=object)
    for i, (a, b) in enumerate(parts):
        res[i, 0], res[i, 1] = a, b
    if der:
        return (res[0], darr[0]) if single else (res, darr)
    return res[0] if single else res


def set_orbit_groups(stride=-1, orbits=-1, min_batch=65536):
    """Group mode of the DOP853 orbit kernel (no counterpart in the reference: a scheduling knob
    of this library).  In batches of at least min_batch orbits, one wavefront in `stride`
    integrates the first `orbits` orbits of the processing order (the longest expected) with 8
    lanes per orbit, each lane evaluating a share of the azimuthal harmonics.  stride=-1: chosen
    from the orbits per lane of the launch; 0: off; orbits=-1: one per 8 lanes of those
    wavefronts.  Results differ from the one-lane form only in the order of the additions of the
    Fourier sum.  The setting is process-wide."""
    check(lib.agx_set_orbit_groups(int(stride), int(orbits), int(min_batch)), "set_orbit_groups")


class Target:
    """A target plug-in collecting data along orbits - reference: agama.Target
    (src/interface_python.cpp:7005-7250).  Provided: type='DensitySphHarm' with gridr (radial
    nodes), lmax, mmax (even orders of the spherical-harmonic expansion): the density
    discretisation of TargetDensitySphHarm (src/galaxymodel_densitygrid.cpp:356-411); and
    type='DensityCylindricalTopHat' / 'DensityCylindricalLinear' with gridr, gridz (nodes in
    R and |z|) and mmax (TargetDensityCylindrical<0/1>, :486-545); type='DensityClassicTopHat' /
    'DensityClassicLinear' with gridr (shell radii), stripsPerPane (default 2), axisRatioY and
    axisRatioZ (default 1) (TargetDensityClassic<0/1>, :146-239); and type='KinemShell' with
    gridr (nodes in r, from r = 0) and degree 0..3: v_r^2 and v_t^2 on B-splines of that degree
    (TargetKinemShell<N>, src/galaxymodel_losvd.cpp:568-576; interface_python.cpp:7228-7243).
    len(target) is the number of values recorded per orbit."""

    def __init__(self, **kwargs):
        self._h = None
        type_str = str(kwargs.pop("type", "") or "")
        if not type_str:
            raise TypeError("Must provide a type='...' argument")
        kind = type_str.lower()
        names = {"densitysphharm": "DensitySphHarm", "densitycylindricaltophat": "DensityCylindricalTopHat",
                 "densitycylindricallinear": "DensityCylindricalLinear", "kinemshell": "KinemShell",
                 "densityclassictophat": "DensityClassicTopHat", "densityclassiclinear": "DensityClassicLinear"}
        if kind not in names:
            raise AgamaAmdError(f"Target type {type_str} is not provided by this version "
                                "(DensitySphHarm, DensityCylindricalTopHat, DensityCylindricalLinear, "
                                "DensityClassicTopHat, DensityClassicLinear, KinemShell are)")
        if "gridr" not in kwargs:
            raise TypeError("Argument 'gridr' is required")
        gridr = np.ascontiguousarray(kwargs.pop("gridr"), dtype=np.float64).reshape(-1)
        if len(gridr) < 2:
            raise ValueError("gridr must be an array with >=2 elements")
        self.type, self.gridr = names[kind], gridr
        if kind == "densitysphharm":
            lmax = int(kwargs.pop("lmax", 0))
            mmax = int(kwargs.pop("mmax", 0))
            self.lmax, self.mmax = lmax, mmax
            if kwargs:
                raise TypeError("Unknown arguments: " + ", ".join(kwargs))
            h = lib.agx_target_density_sphharm(lmax, mmax, ptr(gridr), len(gridr))
        elif kind.startswith("densityclassic"):   # interface_python.cpp:7191-7203
            axisRatioY = float(kwargs.pop("axisRatioY", 1.0))
            axisRatioZ = float(kwargs.pop("axisRatioZ", 1.0))
            strips = int(kwargs.pop("stripsPerPane", 2))
            self.stripsPerPane, self.axisRatioY, self.axisRatioZ = strips, axisRatioY, axisRatioZ
            if kwargs:
                raise TypeError("Unknown arguments: " + ", ".join(kwargs))
            h = lib.agx_target_density_classic(0 if kind.endswith("tophat") else 1, strips, ptr(gridr), len(gridr),
                                               axisRatioY, axisRatioZ)
        elif kind == "kinemshell":
            degree = int(kwargs.pop("degree", -1))
            if not 0 <= degree <= 3:
                raise ValueError("KinemShell: degree of interpolation should be between 0 and 3")
            self.degree = degree
            if kwargs:
                raise TypeError("Unknown arguments: " + ", ".join(kwargs))
            h = lib.agx_target_kinem_shell(degree, ptr(gridr), len(gridr))
        else:   # interface_python.cpp:7209-7221
            if "gridz" not in kwargs:
                raise TypeError("Argument 'gridz' is required")
            gridz = np.ascontiguousarray(kwargs.pop("gridz"), dtype=np.float64).reshape(-1)
            if len(gridz) < 2:
                raise ValueError("gridz must be an array with >=2 elements")
            mmax = int(kwargs.pop("mmax", 0))
            self.gridz, self.mmax = gridz, mmax
            if kwargs:
                raise TypeError("Unknown arguments: " + ", ".join(kwargs))
            h = lib.agx_target_density_cylindrical(0 if kind.endswith("tophat") else 1, mmax, ptr(gridr), len(gridr),
                                                   ptr(gridz), len(gridz))
        if not h:
            raise ValueError(_native.last_error())
        self._h = h

    def __len__(self):
        return lib.agx_target_num_values(self._h)

    def __repr__(self):
        return f"{self.type} target with {len(self)} values"

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and lib is not None:   # not during interpreter shutdown
            lib.agx_target_destroy(h)


from .surrogate import Surrogate  # noqa: E402
from .orbitobj import Orbit  # noqa: E402,F401
