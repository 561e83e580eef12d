"""Scenario description (host side) and its packing into af_scenario_desc.

A Scenario holds what src/main.cpp assembles before its time loop
(main.cpp:50-127, 166-194): the DEM, the physics constants, the four line
boundary flags, the point boundaries in file order and the rain series.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import (AF_BC_FREE, AF_BC_HFIX, AF_BC_HVAR, AF_BC_QFIX, AF_BC_QVAR)

BC_NAMES = {AF_BC_FREE: "FREE", AF_BC_HFIX: "HFIX", AF_BC_HVAR: "HVAR", AF_BC_QFIX: "QFIX",
            AF_BC_QVAR: "QVAR"}


@dataclass
class PointBC:
    """One point boundary (src/types.hpp:58-66).  x, y are grid coordinates;
    QFIX/QVAR values are per unit width (m^2/s, main.cpp:253), HFIX/HVAR
    values are depths (kernels.cpp:270-273)."""
    x: int
    y: int
    type: int
    value: float = 0.0
    times: Optional[np.ndarray] = None
    values: Optional[np.ndarray] = None
    name: str = ""


@dataclass
class Scenario:
    dem: np.ndarray                 # (ny, nx) float64, row 0 = north
    dx: float
    nodata: float = -9999.0
    fpn: float = 0.035
    depth_thresh: float = 0.001
    max_hflow: float = 10.0         # main.cpp:118
    init_tstep: float = 10.0
    cfl: float = 0.7
    g: float = 9.81                 # main.cpp:123
    theta: float = 1.0              # main.cpp:124
    sim_time: float = 3600.0
    mass_int: float = 60.0
    west_free: bool = False
    east_free: bool = False
    north_free: bool = False
    south_free: bool = False
    points: List[PointBC] = field(default_factory=list)
    rain_times: Optional[np.ndarray] = None   # s
    rain_rates: Optional[np.ndarray] = None   # m/s
    h0: Optional[np.ndarray] = None
    name: str = ""

    @property
    def ny(self):
        return int(self.dem.shape[0])

    @property
    def nx(self):
        return int(self.dem.shape[1])

    @property
    def cells(self):
        return self.nx * self.ny

    def pack(self):
        """Returns a PackedScenario whose `.desc` is an af_scenario_desc."""
        return PackedScenario(self)


class PackedScenario:
    """Flat numpy arrays + the ctypes descriptor pointing at them."""

    def __init__(self, s: Scenario):
        self.dem = np.ascontiguousarray(s.dem, dtype=np.float64)
        n = len(s.points)
        self.px = np.array([p.x for p in s.points], dtype=np.int32) if n else np.zeros(1, np.int32)
        self.py = np.array([p.y for p in s.points], dtype=np.int32) if n else np.zeros(1, np.int32)
        self.pt = np.array([p.type for p in s.points], dtype=np.int32) if n else np.zeros(1, np.int32)
        self.pv = np.array([p.value for p in s.points], dtype=np.float64) if n else np.zeros(1)
        offs, lens, ts, vs = [], [], [], []
        k = 0
        for p in s.points:
            m = 0 if p.times is None else len(p.times)
            offs.append(k)
            lens.append(m)
            if m:
                ts.append(np.asarray(p.times, dtype=np.float64))
                vs.append(np.asarray(p.values, dtype=np.float64))
            k += m
        self.poff = np.array(offs or [0], dtype=np.int32)
        self.plen = np.array(lens or [0], dtype=np.int32)
        self.st = np.concatenate(ts) if ts else np.zeros(1)
        self.sv = np.concatenate(vs) if vs else np.zeros(1)
        rain = s.rain_times is not None and len(s.rain_times) > 0
        self.rt = np.ascontiguousarray(s.rain_times, dtype=np.float64) if rain else np.zeros(1)
        self.rr = np.ascontiguousarray(s.rain_rates, dtype=np.float64) if rain else np.zeros(1)
        self.h0 = None if s.h0 is None else np.ascontiguousarray(s.h0, dtype=np.float64)
        d = _lib.ScenarioDesc()
        d.nx, d.ny = s.nx, s.ny
        d.dx, d.nodata = float(s.dx), float(s.nodata)
        d.dem = _lib.dptr(self.dem)
        d.fpn, d.depth_thresh, d.max_hflow = s.fpn, s.depth_thresh, s.max_hflow
        d.init_tstep, d.cfl, d.g, d.theta = s.init_tstep, s.cfl, s.g, s.theta
        d.sim_time, d.mass_int = s.sim_time, s.mass_int
        d.west_free, d.east_free = int(s.west_free), int(s.east_free)
        d.north_free, d.south_free = int(s.north_free), int(s.south_free)
        d.n_points = n
        d.pt_x, d.pt_y = _lib.iptr(self.px), _lib.iptr(self.py)
        d.pt_type, d.pt_value = _lib.iptr(self.pt), _lib.dptr(self.pv)
        d.pt_series_off, d.pt_series_len = _lib.iptr(self.poff), _lib.iptr(self.plen)
        d.n_series = int(k)
        d.series_t, d.series_v = _lib.dptr(self.st), _lib.dptr(self.sv)
        d.rain_enabled = 1 if rain else 0
        d.n_rain = len(self.rt) if rain else 0
        d.rain_t, d.rain_r = _lib.dptr(self.rt), _lib.dptr(self.rr)
        d.h0 = _lib.dptr(self.h0) if self.h0 is not None else None
        self.desc = d
        self.n_series = k
