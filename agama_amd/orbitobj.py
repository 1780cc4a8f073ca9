"""Interpolated orbits: agama.Orbit objects returned by orbit(..., dtype=object).

Reference: the Orbit class of src/interface_python.cpp:6605-6880 (three quintic splines x(t),
y(t), z(t) in the inertial frame built from the recorded trajectory; calling the object
returns the phase-space point rotated back into the frame rotating at Omega, NaN outside the
recorded time span; its length is the number of spline nodes and [i] gives their times).

The splines live in device memory together with the recorded rows (agx_orbit_record in the
C library): they are built there for the whole batch at once on first use, and every call
evaluates on the device.  All Orbit objects of one batch share the record, which is freed
with the last of them."""
import numpy as np

from ._native import AgamaAmdError, check, last_error, lib, ptr


class _Record:
    """owner of an agx_orbit_record handle"""

    def __init__(self, handle, n, der):
        if not handle:
            raise AgamaAmdError(f"orbit failed: {last_error()}")
        self._h = handle
        self.n = n
        self.der = der
        self._info = None

    def info(self):
        """(nodes per series, reversed flag per series); builds the splines on first use"""
        if self._info is None:
            ns = lib.agx_orbit_record_num_series(self._h)
            nodes = np.empty(ns, dtype=np.int64)
            rev = np.empty(ns, dtype=np.int32)
            check(lib.agx_orbit_record_series_info(self._h, ptr(nodes), ptr(rev)), "orbit interpolation")
            self._info = (nodes, rev.astype(bool))
        return self._info

    def fetch(self, rows, der):
        total = int(np.sum(rows))
        times = np.empty(total)
        traj = np.empty((total, 6))
        dev = np.empty(total * 36) if der else None
        check(lib.agx_orbit_record_fetch(self._h, ptr(times), ptr(traj), ptr(dev)), "orbit")
        return times, traj, dev

    def interp(self, series, t, mode):
        t = np.ascontiguousarray(t, dtype=np.float64)
        m = t.shape[0]
        which = np.full(m, series, dtype=np.int64)
        out = np.empty((m, 6 if mode == 0 else 9))
        check(lib.agx_orbit_record_interp(self._h, m, ptr(which), ptr(t), int(mode), ptr(out)),
              "orbit interpolation")
        return out

    def node_times(self, series, count):
        out = np.empty(int(count))
        if count:
            check(lib.agx_orbit_record_node_times(self._h, int(series), ptr(out)), "orbit interpolation")
        return out

    def nodes(self, series):
        """the series' spline nodes [k, 10]: t, x, y, z, x', y', z', x'', y'', z''"""
        k = int(self.info()[0][series])
        out = np.empty((k, 10))
        if k:
            check(lib.agx_orbit_record_nodes(self._h, int(series), ptr(out)), "orbit interpolation")
        return out

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and lib is not None:   # not during interpreter shutdown
            lib.agx_orbit_record_free(h)


def _time_argument(t):
    """a scalar (Python float or int, as the reference's PyFloat/PyInt check) or a 1d array"""
    if isinstance(t, (float, int)) and not isinstance(t, bool):
        return np.array([float(t)]), True
    arr = np.asarray(t, dtype=np.float64)
    if arr.ndim != 1:
        raise TypeError("Input should be a single number or a 1d array")
    return arr, False


class QuinticPieces:
    """A quintic spline given by its nodes t[k] (increasing), values f, first and second
    derivatives d, s2 (the QuinticSpline of src/math_spline.cpp:1391-1429), held as one
    polynomial in u = (t - t_i) / h_i per segment, and continued linearly beyond the end nodes.
    Host-side operations of the reference's Spline objects on it (interface_python.cpp:6159-6330):
    evaluation up to the third derivative, integrals with a monomial or another spline as weight,
    extrema and roots."""

    def __init__(self, t, f, d, s2):
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.f = np.asarray(f, dtype=np.float64)
        self.d = np.asarray(d, dtype=np.float64)
        h = np.diff(self.t)
        df = np.diff(self.f)
        d0, d1 = h * self.d[:-1], h * self.d[1:]
        s0, s1 = h * h * np.asarray(s2)[:-1], h * h * np.asarray(s2)[1:]
        # the quintic with value, derivative and second derivative given at u = 0 and u = 1
        c = np.empty((len(h), 6))
        c[:, 0] = self.f[:-1]
        c[:, 1] = d0
        c[:, 2] = 0.5 * s0
        c[:, 3] = 10 * df - 6 * d0 - 4 * d1 - 1.5 * s0 + 0.5 * s1
        c[:, 4] = -15 * df + 8 * d0 + 7 * d1 + 1.5 * s0 - s1
        c[:, 5] = 6 * df - 3 * d0 - 3 * d1 - 0.5 * s0 + 0.5 * s1
        self.c, self.h = c, h

    def _segment(self, x):
        # binSearch's convention: the last node belongs to the last segment, beyond it is outside
        return np.clip(np.searchsorted(self.t, x, side="right") - 1, 0, len(self.h) - 1)

    def __call__(self, x, der=0):
        x = np.asarray(x, dtype=np.float64)
        out = np.full(x.shape, np.nan)
        if len(self.t) == 0:
            return out
        if len(self.t) == 1:
            inside = np.zeros(x.shape, bool)
        else:
            inside = (x >= self.t[0]) & (x <= self.t[-1])
        if inside.any():
            xi = x[inside]
            i = self._segment(xi)
            u = (xi - self.t[i]) / self.h[i]
            c = self.c[i]
            # Horner for the der-th derivative in u (d^der u^k / du^der = k!/(k-der)! u^(k-der)),
            # then d/dt = (d/du) / h
            v = np.zeros(len(xi))
            for k in range(5, der - 1, -1):
                v = v * u + c[:, k] * float(np.prod(np.arange(k - der + 1, k + 1)))
            out[inside] = v / self.h[i] ** der
        below, above = x < self.t[0], x > self.t[-1]
        for m, j in ((below, 0), (above, -1)):
            if m.any():
                dj = self.d[j]
                out[m] = (self.f[j] + (0 if dj == 0 else dj * (x[m] - self.t[j])) if der == 0 else
                          dj if der == 1 else 0.)
        return out

    def integrate(self, x1, x2, n=0):
        """integral of the spline times x^n, or times another QuinticPieces, over x1..x2; the
        spline is zero outside its grid (integratePiecewise, src/math_spline.cpp:741-843)"""
        sign = 1.
        if x1 > x2:
            x1, x2, sign = x2, x1, -1.
        if len(self.t) < 2:
            return 0.
        other = n if isinstance(n, QuinticPieces) else None
        lo, hi = max(x1, self.t[0]), min(x2, self.t[-1])
        if other is not None:
            if len(other.t) < 2:
                return 0.
            lo, hi = max(lo, other.t[0]), min(hi, other.t[-1])
            order = 6             # a product of two quintics: degree 10
        else:
            order = (5 + int(n)) // 2 + 1
        if not lo < hi:
            return 0.
        grid = self.t if other is None else np.union1d(self.t, other.t)
        edges = np.concatenate([[lo], grid[(grid > lo) & (grid < hi)], [hi]])
        gx, gw = np.polynomial.legendre.leggauss(order)
        a, b = edges[:-1, None], edges[1:, None]
        xs = (0.5 * (b - a) * (gx + 1) + a).ravel()
        w = (0.5 * (b - a) * gw).ravel()
        vals = self(xs) * (other(xs) if other is not None else xs ** int(n))
        return sign * float(np.sum(vals * w))

    def convolve(self, x, kernel):
        """the spline convolved with a Gaussian of width `kernel` (a number > 0) or with another
        QuinticPieces, at the points x: the integral over the spline's grid of S(y) K(x - y) dy,
        the spline being zero outside its grid (integratePiecewise<N, true>,
        src/math_spline.cpp:733-805).  Gaussian: Gauss-Legendre rules of 16 points on pieces no
        wider than sigma/2 within 12 sigma of x (the rest weighs below 1e-31); another spline:
        6-point rules between the breakpoints of both, exact for the degree-10 products."""
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        out = np.zeros(x.shape)
        if len(self.t) < 2:
            return out
        lo, hi = self.t[0], self.t[-1]
        if isinstance(kernel, QuinticPieces):
            gx, gw = np.polynomial.legendre.leggauss(6)
            for j, xj in enumerate(x):
                if len(kernel.t) < 2:
                    continue
                a, b = max(lo, xj - kernel.t[-1]), min(hi, xj - kernel.t[0])
                if not a < b:
                    continue
                br = np.union1d(self.t, xj - kernel.t)
                e = np.concatenate([[a], br[(br > a) & (br < b)], [b]])
                y = (0.5 * (e[1:, None] - e[:-1, None]) * (gx + 1) + e[:-1, None]).ravel()
                w = (0.5 * (e[1:, None] - e[:-1, None]) * gw).ravel()
                out[j] = np.sum(w * self(y) * kernel(xj - y))
            return out
        sigma = float(kernel)
        if not sigma > 0:
            raise ValueError("the Gaussian width must be positive")
        gx, gw = np.polynomial.legendre.leggauss(16)
        norm = 1 / (np.sqrt(2 * np.pi) * sigma)
        for j, xj in enumerate(x):
            a, b = max(lo, xj - 12 * sigma), min(hi, xj + 12 * sigma)
            if not a < b:
                continue
            npieces = int(np.ceil((b - a) / (0.5 * sigma)))
            br = np.union1d(np.linspace(a, b, npieces + 1), self.t[(self.t > a) & (self.t < b)])
            y = (0.5 * (br[1:, None] - br[:-1, None]) * (gx + 1) + br[:-1, None]).ravel()
            w = (0.5 * (br[1:, None] - br[:-1, None]) * gw).ravel()
            out[j] = norm * np.sum(w * self(y) * np.exp(-0.5 * ((xj - y) / sigma) ** 2))
        return out

    def extrema(self, x1=np.nan, x2=np.nan):
        """points of local minimum or maximum on x1..x2 (default: the grid) in increasing order,
        with both ends of the interval (findRoots<true>, src/math_spline.cpp:915-975)"""
        if len(self.t) == 0:
            return np.empty(0)
        if x1 > x2:
            x1, x2 = x2, x1
        x1 = self.t[0] if not x1 >= self.t[0] else x1
        x2 = self.t[-1] if not x2 <= self.t[-1] else x2
        res, prev, sign = [x1], float(self(np.array([x1]))[0]), 0.
        if x1 == x2 or len(self.t) < 2:
            return np.array(res)
        i1, i2 = int(self._segment(np.array([x1]))[0]), int(self._segment(np.array([x2]))[0])
        for i in range(i1, i2 + 1):
            t1, t2 = max(x1, self.t[i]), min(x2, self.t[i + 1])
            c = self.c[i]
            # zeros of dp/du = c1 + 2 c2 u + 3 c3 u^2 + 4 c4 u^3 + 5 c5 u^4
            dp = np.array([5 * c[5], 4 * c[4], 3 * c[3], 2 * c[2], c[1]])
            nz = np.flatnonzero(dp)
            cand = []
            if len(nz) and nz[0] < 4:
                r = np.roots(dp[nz[0]:])
                cand = sorted(self.t[i] + self.h[i] * r.real[np.abs(r.imag) <= 1e-12 * np.maximum(1, np.abs(r.real))])
            if i == i2:
                cand.append(t2)
            for xr in cand:
                if t1 <= xr <= t2 and xr > res[-1] + 10 * np.finfo(float).eps * max(abs(xr), abs(res[-1])):
                    v = float(self(np.array([xr]))[0])
                    if (sign > 0 and v > prev) or (sign < 0 and v < prev):
                        res[-1], prev = xr, v      # the trend continues: an inflection, not an extremum
                    elif v != prev:
                        sign = 1. if v > prev else -1.
                        res.append(xr)
                        prev = v
        return np.array(res)

    def roots(self, y=0., x1=np.nan, x2=np.nan):
        """all points of x1..x2 where the spline equals y, in increasing order: one root on each
        interval between consecutive extrema where the sign of spline - y changes
        (QuinticSpline::roots, src/math_spline.cpp:1446-1461)"""
        from scipy.optimize import brentq
        ex = self.extrema(x1, x2)
        res = []
        for a, b in zip(ex[:-1], ex[1:]):
            fa, fb = float(self(np.array([a]))[0]) - y, float(self(np.array([b]))[0]) - y
            if fa == 0:
                r = a
            elif fb == 0:
                r = b
            elif fa * fb < 0:
                r = brentq(lambda q: float(self(np.array([q]))[0]) - y, a, b, xtol=1e-15 * max(abs(a), abs(b), 1.),
                           rtol=4 * np.finfo(float).eps)
            else:
                continue
            if not res or r > res[-1] + 10 * np.finfo(float).eps * max(abs(r), abs(res[-1])):
                res.append(r)
        return np.array(res)


class OrbitSpline:
    """one coordinate (x, y or z, inertial frame) of an interpolated orbit as a quintic spline
    through the recorded positions with the velocities as derivatives: spline(t, der=0) with
    der = 0..3, ext (a value for points outside the recorded span instead of the linear
    continuation beyond the end nodes), conv (a Gaussian width or another spline to convolve
    with), and the reference's Spline methods integrate, roots and extrema (QuinticSpline, src/math_spline.cpp:1391-1463; Spline_value / Spline_integrate /
    Spline_roots / Spline_extrema, src/interface_python.cpp:6159-6330).  der 0-2 are evaluated on
    the device, the third derivative, convolutions and the methods on the host from the spline's
    nodes."""

    def __init__(self, rec, series, comp):
        self._rec, self._series, self._comp = rec, series, comp
        self._pieces = None

    def _host(self):
        if self._pieces is None:
            nd = self._rec.nodes(self._series)
            k = self._comp
            self._pieces = QuinticPieces(nd[:, 0], nd[:, 1 + k], nd[:, 4 + k], nd[:, 7 + k])
        return self._pieces

    def __call__(self, x, der=0, ext=None, conv=None):
        if der not in (0, 1, 2, 3):
            raise ValueError("Can only compute derivatives up to 3rd")
        if conv is not None and (der != 0 or ext is not None):
            raise ValueError("`conv` cannot be used together with `der` or `ext`")
        t, scalar = _time_argument(x)
        if conv is not None:
            out = self._convolve(t, conv)
            return float(out[0]) if scalar else out
        if der == 3:
            out = self._host()(t, 3)
        else:
            out = self._rec.interp(self._series, t, 1)[:, der * 3 + self._comp]
        if ext is not None:
            nodes = self._host().t
            if len(nodes):
                out = np.where((t < nodes[0]) | (t > nodes[-1]), float(ext), out)
        return float(out[0]) if scalar else out

    def _convolve(self, t, conv):
        # conv: another Orbit spline, or Gaussian widths (a number, or one per point); a width
        # of zero gives the spline's value inside its grid and 0 outside (Spline_value,
        # src/interface_python.cpp:6257-6270)
        host = self._host()
        if isinstance(conv, OrbitSpline):
            return host.convolve(t, conv._host())
        sig = np.asarray(conv, dtype=np.float64)
        if sig.size != 1 and sig.shape != t.shape:
            raise TypeError("Argument `sigma`, if provided, must be a single number or an array with the "
                            "same shape as `x`")
        sig = np.broadcast_to(sig.reshape(-1) if sig.size != 1 else sig.reshape(()), t.shape)
        out = np.empty(t.shape)
        for j in range(len(t)):
            if sig[j] == 0:
                inside = len(host.t) and host.t[0] <= t[j] <= host.t[-1]
                out[j] = float(host(np.array([t[j]]))[0]) if inside else 0.
            else:
                out[j] = host.convolve(t[j:j + 1], float(sig[j]))[0]
        return out

    def integrate(self, x1, x2, n=0):
        """integral over x1..x2 of the spline times x^n (n a non-negative integer) or times
        another spline of an Orbit; the spline is zero outside its grid"""
        if isinstance(n, OrbitSpline):
            return self._host().integrate(float(x1), float(x2), n._host())
        if isinstance(n, (int, np.integer)) and 0 <= n < 0x7fff:
            return self._host().integrate(float(x1), float(x2), int(n))
        raise RuntimeError("Argument 'n', if present, should be a nonnegative integer or another "
                           "instance of Spline")

    def roots(self, y=0., x1=np.nan, x2=np.nan):
        """all points on x1..x2 (default: the grid) where the spline equals y, in increasing order"""
        return self._host().roots(float(y), float(x1), float(x2))

    def extrema(self, x1=np.nan, x2=np.nan):
        """all local minima and maxima on x1..x2 (default: the grid) with both ends, increasing"""
        return self._host().extrema(float(x1), float(x2))

    def __len__(self):
        return int(self._rec.info()[0][self._series])

    def __repr__(self):
        return "Quintic spline for " + "xyz"[self._comp] + "(t)"


class Orbit:
    """Result of orbit integration represented as three spline interpolators.

    Returned by orbit() with dtype=object.  orbit(t) gives the 6 phase-space coordinates at time
    t (or an N x 6 array for a 1d array of times), NaN outside the recorded time span; len(orbit)
    is the number of recorded points and orbit[i] the time of the i-th one (in the order of
    integration).  Members x, y, z (the inertial-frame splines), Omega (the pattern speed of
    the rotating frame) and reversed (integrated backward in time)."""

    __slots__ = ("_rec", "_series", "Omega", "x", "y", "z", "_times")

    def __init__(self, *args, **kwargs):
        raise TypeError("Instances of Orbit class cannot be created from Python")

    @classmethod
    def _make(cls, rec, series, Omega):
        self = object.__new__(cls)
        self._rec = rec
        self._series = int(series)
        self.Omega = float(Omega)
        self.x = OrbitSpline(rec, series, 0)
        self.y = OrbitSpline(rec, series, 1)
        self.z = OrbitSpline(rec, series, 2)
        self._times = None
        return self

    @property
    def reversed(self):
        return bool(self._rec.info()[1][self._series])

    def __len__(self):
        return int(self._rec.info()[0][self._series])

    def _node_times(self):
        if self._times is None:
            self._times = self._rec.node_times(self._series, len(self))
        return self._times

    def __getitem__(self, index):
        if not isinstance(index, (int, np.integer)):
            raise TypeError("Orbit indices must be integers")
        size = len(self)
        i = int(index) + (size if index < 0 else 0)
        if not 0 <= i < size:
            raise IndexError("Orbit timestamp index out of range")
        return float(self._node_times()[size - 1 - i if self.reversed else i])

    def __call__(self, t):
        arr, scalar = _time_argument(t)
        out = self._rec.interp(self._series, arr, 0)
        return out[0] if scalar else out

    def __repr__(self):
        size = len(self)
        s = f"Orbit with {size} nodes"
        if size:
            nt = self._node_times()
            s += f" on t=[{nt[0]:.16g}:{nt[-1]:.16g}]"
        return s
