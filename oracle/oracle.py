"""ctypes binding of oracle.c, the FP64 CPU restatement of the reference step.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Scenarios are taken
duck-typed (attributes dem, dx, nodata, fpn, ..., points, rain_times,
rain_rates, h0) so that this module imports nothing from the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
# the same oracle with the C library's pow(hflow, 10.0/3.0) in the flux
LIB_LIBM = os.path.join(HERE, "_build", "liboracle_libm.so")
# the same oracle summing the depth its clamps at zero add (orc_clamp_gain)
LIB_CLAMP = os.path.join(HERE, "_build", "liboracle_clamp.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

QFIX, QVAR, FREE, HFIX, HVAR = 4, 5, 1, 2, 3


class OrcCase(C.Structure):
    _fields_ = [
        ("xsz", C.c_int), ("ysz", C.c_int), ("dx", C.c_double), ("nodata", C.c_double),
        ("fpn", C.c_double), ("depth_thresh", C.c_double), ("max_hflow", C.c_double),
        ("init_tstep", C.c_double), ("cfl", C.c_double), ("g", C.c_double),
        ("theta", C.c_double), ("sim_time", C.c_double), ("mass_int", C.c_double),
        ("west_free", C.c_int), ("east_free", C.c_int), ("north_free", C.c_int),
        ("south_free", C.c_int),
        ("n_flow", C.c_int), ("flow_idx", _ip), ("flow_type", _ip), ("flow_value", _dp),
        ("flow_off", _ip), ("flow_len", _ip),
        ("n_other", C.c_int), ("other_idx", _ip), ("other_code", _ip), ("other_value", _dp),
        ("other_off", _ip), ("other_len", _ip),
        ("series_t", _dp), ("series_v", _dp),
        ("rain_enabled", C.c_int), ("n_rain", C.c_int), ("rain_t", _dp), ("rain_r", _dp),
    ]


class OrcRunInfo(C.Structure):
    _fields_ = [
        ("max_steps", C.c_long), ("steps_done", C.c_long), ("t", C.c_double),
        ("dt_log", _dp), ("dt_log_cap", C.c_long), ("mass_log", _dp), ("mass_log_cap", C.c_long),
        ("n_mass", C.c_long), ("step_seconds", C.c_double), ("next_mass", C.c_double),
    ]


_libs = {}


def build(force=False):
    if force or not all(os.path.exists(p) for p in (LIB, LIB_LIBM, LIB_CLAMP)):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-f",
                        os.path.join(HERE, "Makefile")], check=True)
    return LIB


def lib(libm=False, clamp=False):
    """The pinned oracle; with libm=True the build whose flux takes the C
    library's pow (for measuring what that rounding does to a run); with
    clamp=True the build that sums the depth its clamps add (same results)."""
    key = "clamp" if clamp else ("libm" if libm else "pinned")
    if key not in _libs:
        build()
        L = C.CDLL({"clamp": LIB_CLAMP, "libm": LIB_LIBM, "pinned": LIB}[key])
        V = C.c_void_p
        D = C.c_double
        I = C.c_int
        L.orc_update_qs.argtypes = [V, V, V, V, I, I, D]
        L.orc_find_hmax.argtypes = [V, V, C.c_long, D]
        L.orc_find_hmax.restype = D
        L.orc_dt_from_hmax.argtypes = [D, D, D, D, D, D]
        L.orc_dt_from_hmax.restype = D
        L.orc_face_flux.argtypes = [D] * 15
        L.orc_face_flux.restype = D
        L.orc_pow_10_3.argtypes = [D]
        L.orc_pow_10_3.restype = D
        L.orc_floodplain_q.argtypes = [V, V, V, V, V, V, I, I, D, D, D, D, D, D, _dp, D, D, D, D,
                                       D, D, D]
        L.orc_bcs.argtypes = [V, V, V, V, I, I, D, D, D, D, I, I, I, I]
        L.orc_apply_point_bcs.argtypes = [V, V, V, V, I]
        L.orc_update_h.argtypes = [V, V, V, I, I, D, D]
        L.orc_rainfall.argtypes = [V, V, I, I, D, D, D]
        L.orc_add_flow_bcs.argtypes = [V, V, V, I]
        L.orc_interp.argtypes = [V, V, I, D, D]
        L.orc_interp.restype = D
        L.orc_velocity.argtypes = [V, V, V, V, V, V, I, I, D, D]
        L.orc_run.argtypes = [C.POINTER(OrcCase), V, V, V, V, V, V, D, D, C.POINTER(OrcRunInfo)]
        L.orc_run.restype = I
        L.orc_num_threads.restype = I
        L.orc_set_num_threads.argtypes = [I]
        L.orc_clamp_gain.argtypes = [I]
        L.orc_clamp_gain.restype = D
        _libs[key] = L
    return _libs[key]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


class _Packed:
    def __init__(self, s):
        self.dem = np.ascontiguousarray(s.dem, dtype=np.float64).ravel()
        nx = s.dem.shape[1]
        flow = [p for p in s.points if p.type in (QFIX, QVAR)]
        other = [p for p in s.points if p.type in (FREE, HFIX, HVAR)]
        series_t, series_v = [], []

        def ser(p):
            off = sum(len(a) for a in series_t)
            n = 0 if p.times is None else len(p.times)
            if n:
                series_t.append(np.asarray(p.times, dtype=np.float64))
                series_v.append(np.asarray(p.values, dtype=np.float64))
            return off, n

        fo = [ser(p) for p in flow]
        oo = [ser(p) for p in other]
        code = {FREE: 1, HFIX: 2, HVAR: 3}
        i32 = lambda v: np.array(v or [0], dtype=np.int32)  # noqa: E731
        self.flow_idx = i32([p.x + p.y * nx for p in flow])
        self.flow_type = i32([p.type for p in flow])
        self.flow_value = np.array([p.value for p in flow] or [0.0])
        self.flow_off = i32([o for o, _ in fo])
        self.flow_len = i32([n for _, n in fo])
        self.other_idx = i32([p.x + p.y * nx for p in other])
        self.other_code = i32([code[p.type] for p in other])
        self.other_value = np.array([p.value for p in other] or [0.0])
        self.other_off = i32([o for o, _ in oo])
        self.other_len = i32([n for _, n in oo])
        self.st = np.concatenate(series_t) if series_t else np.zeros(1)
        self.sv = np.concatenate(series_v) if series_v else np.zeros(1)
        rain = s.rain_times is not None and len(s.rain_times) > 0
        self.rt = np.ascontiguousarray(s.rain_times, dtype=np.float64) if rain else np.zeros(1)
        self.rr = np.ascontiguousarray(s.rain_rates, dtype=np.float64) if rain else np.zeros(1)
        c = OrcCase()
        c.xsz, c.ysz = s.dem.shape[1], s.dem.shape[0]
        c.dx, c.nodata = float(s.dx), float(s.nodata)
        c.fpn, c.depth_thresh, c.max_hflow = s.fpn, s.depth_thresh, s.max_hflow
        c.init_tstep, c.cfl, c.g, c.theta = s.init_tstep, s.cfl, s.g, s.theta
        c.sim_time, c.mass_int = s.sim_time, s.mass_int
        c.west_free, c.east_free = int(s.west_free), int(s.east_free)
        c.north_free, c.south_free = int(s.north_free), int(s.south_free)
        c.n_flow = len(flow)
        c.flow_idx, c.flow_type = self.flow_idx.ctypes.data_as(_ip), self.flow_type.ctypes.data_as(_ip)
        c.flow_value = self.flow_value.ctypes.data_as(_dp)
        c.flow_off, c.flow_len = self.flow_off.ctypes.data_as(_ip), self.flow_len.ctypes.data_as(_ip)
        c.n_other = len(other)
        c.other_idx = self.other_idx.ctypes.data_as(_ip)
        c.other_code = self.other_code.ctypes.data_as(_ip)
        c.other_value = self.other_value.ctypes.data_as(_dp)
        c.other_off, c.other_len = self.other_off.ctypes.data_as(_ip), self.other_len.ctypes.data_as(_ip)
        c.series_t, c.series_v = self.st.ctypes.data_as(_dp), self.sv.ctypes.data_as(_dp)
        c.rain_enabled = 1 if rain else 0
        c.n_rain = len(self.rt) if rain else 0
        c.rain_t, c.rain_r = self.rt.ctypes.data_as(_dp), self.rr.ctypes.data_as(_dp)
        self.case = c


class OracleRun:
    """State of one oracle run; step it with advance()."""

    def __init__(self, s, dt_log_cap=1 << 20, mass_log_cap=4096, libm=False, clamp=False):
        self.s = s
        self.libm = libm
        self.clamp = clamp
        self.pk = _Packed(s)
        ny, nx = s.dem.shape
        self.H = np.zeros(nx * ny)
        if getattr(s, "h0", None) is not None:
            self.H[:] = np.asarray(s.h0, dtype=np.float64).ravel()
        nq = (nx + 1) * (ny + 1)
        self.Qx, self.Qy, self.Qxold, self.Qyold = (np.zeros(nq) for _ in range(4))
        self.t = 0.0
        self.next_mass = float(s.mass_int)
        self.steps = 0
        self.dts = []
        self.reports = []
        self.step_seconds = 0.0
        self._dt_cap = dt_log_cap
        self._mass_cap = mass_log_cap

    def advance(self, max_steps=None):
        """Runs until t >= sim_time or max_steps more steps."""
        info = OrcRunInfo()
        info.max_steps = int(max_steps) if max_steps else 0
        dtl = np.zeros(max(self._dt_cap, 1))
        ml = np.zeros(4 * self._mass_cap)
        info.dt_log, info.dt_log_cap = dtl.ctypes.data_as(_dp), self._dt_cap
        info.mass_log, info.mass_log_cap = ml.ctypes.data_as(_dp), self._mass_cap
        L = lib(self.libm, self.clamp)
        rc = L.orc_run(C.byref(self.pk.case), _p(self.H), _p(self.pk.dem), _p(self.Qx),
                       _p(self.Qy), _p(self.Qxold), _p(self.Qyold), self.t, self.next_mass,
                       C.byref(info))
        if rc != 0:
            raise RuntimeError("oracle run failed")
        n = int(info.steps_done)
        self.dts.extend(dtl[: min(n, self._dt_cap)].tolist())
        nm = int(info.n_mass)
        self.reports.extend(ml[: 4 * min(nm, self._mass_cap)].reshape(-1, 4).tolist())
        self.next_mass = info.next_mass
        self.t = info.t
        self.steps += n
        self.step_seconds += info.step_seconds
        return n

    def state(self):
        ny, nx = self.s.dem.shape
        return {"H": self.H.reshape(ny, nx), "Qx": self.Qx, "Qy": self.Qy, "Qxold": self.Qxold,
                "Qyold": self.Qyold, "t": self.t, "steps": self.steps, "dt": np.array(self.dts),
                "reports": np.array(self.reports).reshape(-1, 4)}


def run(s, max_steps=None, libm=False):
    r = OracleRun(s, libm=libm)
    r.advance(max_steps)
    return r.state()


# per-function entry points, for known-answer tests ------------------------

def floodplain_q(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, fpn, thresh, max_hflow, cfl, g,
                 init_tstep, theta, nodata):
    ts = C.c_double(0.0)
    lib().orc_floodplain_q(_p(H), _p(DEM), _p(Qx), _p(Qy), _p(Qxold), _p(Qyold), xsz, ysz, dx, dx,
                           fpn, thresh, max_hflow, 0.01, C.byref(ts), 1.0, dx * dx, cfl, g,
                           init_tstep, theta, nodata)
    return ts.value


def bcs(Qx, Qy, H, DEM, xsz, ysz, dx, g, nodata, w, e, n, s):
    lib().orc_bcs(_p(Qx), _p(Qy), _p(H), _p(DEM), xsz, ysz, dx, g, 0.0, nodata, int(w), int(e),
                  int(n), int(s))


def update_h(H, Qx, Qy, xsz, ysz, dA, Tstep):
    lib().orc_update_h(_p(H), _p(Qx), _p(Qy), xsz, ysz, dA, Tstep)


def rainfall(H, DEM, xsz, ysz, rate, Tstep, nodata):
    lib().orc_rainfall(_p(H), _p(DEM), xsz, ysz, rate, Tstep, nodata)


def add_flow_bcs(H, idx, ps):
    lib().orc_add_flow_bcs(_p(H), _p(idx), _p(ps), len(idx))


def apply_point_bcs(H, idx, code, value):
    lib().orc_apply_point_bcs(_p(H), _p(idx), _p(code), _p(value), len(idx))


def update_qs(Qx, Qy, Qxold, Qyold, xsz, ysz, dx):
    lib().orc_update_qs(_p(Qx), _p(Qy), _p(Qxold), _p(Qyold), xsz, ysz, dx)


def interp(times, values, empty_value, t):
    ts = np.ascontiguousarray(times, dtype=np.float64)
    vs = np.ascontiguousarray(values, dtype=np.float64)
    return lib().orc_interp(_p(ts), _p(vs), len(ts), float(empty_value), float(t))


def face_flux(z0, z1, h0, h1, q0, qup, qdown, qt, dx, fn, thresh, maxh, dt, g, theta):
    return lib().orc_face_flux(z0, z1, h0, h1, q0, qup, qdown, qt, dx, fn, thresh, maxh, dt, g,
                               theta)


def pow_10_3(h):
    """The oracle's pow(h, 10.0/3.0) (IEEE-exact operations only)."""
    return lib().orc_pow_10_3(float(h))


def find_hmax(H, DEM, nodata):
    return lib().orc_find_hmax(_p(H), _p(DEM), len(H), nodata)


def velocity(H, DEM, Qx, Qy, xsz, ysz, dx, nodata):
    Vx = np.zeros(xsz * ysz)
    Vy = np.zeros(xsz * ysz)
    lib().orc_velocity(_p(Vx), _p(Vy), _p(H), _p(DEM), _p(Qx), _p(Qy), xsz, ysz, dx, nodata)
    return Vx, Vy
