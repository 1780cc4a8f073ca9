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


class OrbitSpline:
    """one coordinate (x, y or z, inertial frame) of an interpolated orbit as a quintic spline
    through the recorded positions with the velocities as derivatives: spline(t, der=0) with
    der = 0, 1, 2; continued linearly beyond the end nodes (QuinticSpline::evalDeriv,
    src/math_spline.cpp:1405-1429)"""

    def __init__(self, rec, series, comp):
        self._rec, self._series, self._comp = rec, series, comp

    def __call__(self, x, der=0):
        if der not in (0, 1, 2):
            raise ValueError("der must be 0, 1 or 2")
        t, scalar = _time_argument(x)
        out = self._rec.interp(self._series, t, 1)[:, der * 3 + self._comp]
        return float(out[0]) if scalar else out

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
