"""Input files of the reference (.par, .dem.ascii, .bci, .bdy, rain) and its
ESRI-ASCII writer, through the C++ readers of libadaptflood_amd.so.

load_case(par_path) runs the reference's input sequence (src/main.cpp:24-92)
and returns a Scenario ready for the step.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import Par, call
from .scenario import PointBC, Scenario


def _b(s):
    return os.fsencode(s) if isinstance(s, str) else s


def _par_dict(p):
    d = {k: getattr(p, k) for k, _ in Par._fields_}
    for k in ("dem_file", "bci_file", "bdy_file", "rain_file", "stage_file", "output_dir"):
        d[k] = d[k].decode()
    return d


def par_defaults():
    p = Par()
    _lib.lib().af_par_defaults(C.byref(p))
    return _par_dict(p)


def read_par(path):
    """src/io.cpp:9-47: keys as the code reads them."""
    p = Par()
    _lib.lib().af_par_defaults(C.byref(p))
    call("af_read_par", _b(path), C.byref(p))
    return _par_dict(p)


class Case:
    """A parsed case (DEM, points, line boundaries, series, rain)."""

    def __init__(self):
        self.h = _lib.lib().af_case_new()

    def __del__(self):
        if getattr(self, "h", None):
            _lib.lib().af_case_free(self.h)
            self.h = None

    def read_dem(self, path):
        call("af_case_read_dem", self.h, _b(path))

    def read_bci(self, path):
        call("af_case_read_bci", self.h, _b(path))

    def read_bdy(self, path):
        call("af_case_read_bdy", self.h, _b(path))

    def read_rain(self, path):
        call("af_case_read_rain", self.h, _b(path))

    def load(self, par_path):
        call("af_case_load", self.h, _b(par_path))

    def par(self):
        p = Par()
        call("af_case_par", self.h, C.byref(p))
        return _par_dict(p)

    def header(self):
        nc, nr = C.c_int32(), C.c_int32()
        xll, yll, cs, nod = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        call("af_case_dem_header", self.h, C.byref(nc), C.byref(nr), C.byref(xll), C.byref(yll),
             C.byref(cs), C.byref(nod))
        return {"ncols": nc.value, "nrows": nr.value, "xll": xll.value, "yll": yll.value,
                "cellsize": cs.value, "nodata": nod.value}

    def line_bcs(self):
        n = C.c_int32()
        call("af_case_line_bcs", self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(max(n.value, 1))
        call("af_case_line_bcs", self.h, buf, n.value, C.byref(n))
        return buf.raw[: n.value].decode()

    def points(self):
        out = []
        i = 0
        while True:
            x, y, ty, ln = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            v = C.c_double()
            name = C.create_string_buffer(128)
            rc = _lib.lib().af_case_point(self.h, i, C.byref(x), C.byref(y), C.byref(ty),
                                          C.byref(v), C.byref(ln), name, 128)
            if rc != 0:
                break
            times = values = None
            if ln.value:
                times = np.empty(ln.value)
                values = np.empty(ln.value)
                call("af_case_point_series", self.h, i, _lib.dptr(times), _lib.dptr(values),
                     ln.value)
            out.append(PointBC(x.value, y.value, ty.value, v.value, times, values,
                               name.value.decode()))
            i += 1
        return out

    def rain(self):
        en, n = C.c_int32(), C.c_int32()
        call("af_case_rain", self.h, C.byref(en), None, None, 0, C.byref(n))
        t = np.empty(max(n.value, 1))
        r = np.empty(max(n.value, 1))
        call("af_case_rain", self.h, C.byref(en), _lib.dptr(t), _lib.dptr(r), n.value,
             C.byref(n))
        return bool(en.value), t[: n.value], r[: n.value]

    def read_stage(self, path):
        call("af_case_read_stage", self.h, path.encode())

    def stages(self):
        """Stage gauges as (name, grid_x, grid_y) in file order."""
        n = C.c_int32()
        call("af_case_stage", self.h, -1, None, None, None, 0, C.byref(n))
        out = []
        for i in range(n.value):
            gx, gy = C.c_int32(), C.c_int32()
            name = C.create_string_buffer(256)
            call("af_case_stage", self.h, i, C.byref(gx), C.byref(gy), name, 256, C.byref(n))
            out.append((name.value.decode(), gx.value, gy.value))
        return out

    def scenario(self):
        """The Scenario main.cpp would step (constants of main.cpp:115-127)."""
        d = _lib.ScenarioDesc()
        call("af_case_scenario", self.h, C.byref(d))
        nx, ny = d.nx, d.ny
        dem = np.ctypeslib.as_array(d.dem, shape=(ny * nx,)).reshape(ny, nx).copy()
        en, rt, rr = self.rain()
        return Scenario(dem=dem, dx=d.dx, nodata=d.nodata, fpn=d.fpn,
                        depth_thresh=d.depth_thresh, max_hflow=d.max_hflow,
                        init_tstep=d.init_tstep, cfl=d.cfl, g=d.g, theta=d.theta,
                        sim_time=d.sim_time, mass_int=d.mass_int,
                        west_free=bool(d.west_free), east_free=bool(d.east_free),
                        north_free=bool(d.north_free), south_free=bool(d.south_free),
                        points=self.points(), rain_times=rt if en else None,
                        rain_rates=rr if en else None)


def load_case(par_path):
    c = Case()
    c.load(par_path)
    return c


def write_asc(path, data, xll=0.0, yll=0.0, cellsize=1.0):
    """src/io.cpp:314-334: fixed header, NODATA_value -9999, values %.4f."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    ny, nx = a.shape
    call("af_write_asc", _b(path), _lib.dptr(a), nx, ny, float(xll), float(yll), float(cellsize))


def write_asc_async(path, data, xll=0.0, yll=0.0, cellsize=1.0, nodata=-9999.0):
    """src/async_output.cpp:158-176: corners %.6f, the grid's NODATA value as
    %.0f, values %.4f."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    ny, nx = a.shape
    call("af_write_asc_async", _b(path), _lib.dptr(a), nx, ny, float(xll), float(yll),
         float(cellsize), float(nodata))


class NetCDFWriter:
    """The asynchronous writer's output.nc (src/async_output.cpp:18-54,
    108-141) in the classic 64-bit-offset NetCDF format: append(t, H[, Vx,
    Vy]) adds one record of float grids."""

    def __init__(self, path, ncols, nrows, xll=0.0, yll=0.0, cellsize=1.0, nodata=-9999.0,
                 velocity=False):
        h = C.c_void_p()
        call("af_nc_create", _b(path), int(ncols), int(nrows), float(xll), float(yll),
             float(cellsize), float(nodata), 1 if velocity else 0, C.byref(h))
        self.handle, self.shape, self.velocity = h, (nrows, ncols), velocity

    def append(self, t, H, Vx=None, Vy=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (H, Vx, Vy) if a is not None]
        if any(a.size != self.shape[0] * self.shape[1] for a in arrs):
            raise ValueError("grid size does not match the file")
        ptrs = [_lib.dptr(a) for a in arrs] + [None] * (3 - len(arrs))
        call("af_nc_append", self.handle, float(t), *ptrs)

    def close(self):
        if self.handle:
            call("af_nc_close", self.handle)
            self.handle = None


def write_case(dirpath, scn, name="case", xll=0.0, yll=0.0, rain_header="rain", extra_par=""):
    """Writes a Scenario in the reference's input formats (the inverse of
    load_case): .par, .dem.ascii, .bci, .bdy and an optional rain file.
    Point grid coordinates become cell-centre world coordinates."""
    os.makedirs(dirpath, exist_ok=True)
    ny, nx = scn.dem.shape
    cs = float(scn.dx)
    dem_p = os.path.join(dirpath, f"{name}.dem.ascii")
    with open(dem_p, "w") as f:
        f.write(f"ncols {nx}\nnrows {ny}\nxllcorner {xll!r}\nyllcorner {yll!r}\n"
                f"cellsize {cs!r}\nNODATA_value {scn.nodata!r}\n")
        for j in range(ny):
            f.write(" ".join(repr(float(v)) for v in scn.dem[j]) + "\n")
    tly = yll + ny * cs
    names = {}
    bci_p = os.path.join(dirpath, f"{name}.bci")
    bdy_p = os.path.join(dirpath, f"{name}.bdy")
    kinds = {1: "FREE", 2: "HFIX", 3: "HVAR", 4: "QFIX", 5: "QVAR"}
    with open(bci_p, "w") as f:
        for side, on in (("N", scn.north_free), ("S", scn.south_free), ("E", scn.east_free),
                         ("W", scn.west_free)):
            if on:
                f.write(f"{side} 0 0 FREE\n")
        for k, p in enumerate(scn.points):
            wx = xll + (p.x + 0.5) * cs
            wy = tly - (p.y + 0.5) * cs
            kind = kinds[p.type]
            if kind in ("QVAR", "HVAR"):
                nm = p.name or f"S{k}"
                names[nm] = p
                f.write(f"P {wx!r} {wy!r} {kind} {nm}\n")
            elif kind == "FREE":
                f.write(f"P {wx!r} {wy!r} FREE\n")
            else:
                f.write(f"P {wx!r} {wy!r} {kind} {p.value!r}\n")
    with open(bdy_p, "w") as f:
        f.write("series written by adaptflood_amd.formats.write_case\n")
        for nm, p in names.items():
            n = 0 if p.times is None else len(p.times)
            f.write(f"{nm}\n{n} seconds\n")
            for t, v in zip(p.times if n else [], p.values if n else []):
                f.write(f"{float(v)!r} {float(t)!r}\n")
    rain_line = ""
    if scn.rain_times is not None and len(scn.rain_times):
        rain_p = os.path.join(dirpath, f"{name}.rain")
        with open(rain_p, "w") as f:
            f.write(rain_header + "\n")
            for t, r in zip(scn.rain_times, scn.rain_rates):
                f.write(f"{float(r)!r} {float(t)!r}\n")
        rain_line = f"rainfile {rain_p}\n"
    par_p = os.path.join(dirpath, f"{name}.par")
    with open(par_p, "w") as f:
        f.write(f"DEMfile {dem_p}\nbcifile {bci_p}\nbdyfile {bdy_p}\n"
                f"dirroot {os.path.join(dirpath, 'results')}\n"
                f"sim_time {scn.sim_time!r}\nmassint {scn.mass_int!r}\n"
                f"initial_tstep {scn.init_tstep!r}\ncfl {scn.cfl!r}\nfpfric {scn.fpn!r}\n"
                f"depth_thresh {scn.depth_thresh!r}\n" + rain_line + extra_par)
    return par_p
