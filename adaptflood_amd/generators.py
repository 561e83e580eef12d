"""Seeded synthetic scenarios for the five benchmark workloads.

  plane_case      256 x 256 at 10 m, eastward plane + smoothed noise, one QFIX
                  source at the centre, FREE east edge (the FP64 CPU-runnable case)
  valley_case     4096 x 4096 at 5 m fractal DEM with a carved valley, 2000 QVAR
                  hydrographs, 20 mm/h rain for 0-1800 s, HFIX 2 m west edge
  roofline_case   16384 x 16384 at 2 m fractal DEM, 10% NODATA blobs, 5000 QFIX
  ensemble        512 ragged scenarios, sides log-uniform 128..2048, aspect 1..4
  arrivals        Poisson arrival times for the served ensemble

Point values follow the reference's units: QFIX/QVAR values are per unit
width (m^2/s; the source adds value*dx*dt m^3, main.cpp:253), so a discharge
of Q m^3/s is written as Q/dx.  HFIX/HVAR values are depths (m).
"""
import math

import numpy as np

from ._lib import AF_BC_HFIX, AF_BC_QFIX, AF_BC_QVAR
from .scenario import PointBC, Scenario


def box_smooth(a, k):
    """k x k moving average with edge replication (integral image)."""
    if k <= 1:
        return a.copy()
    p = k // 2
    b = np.pad(a, ((p, k - 1 - p), (p, k - 1 - p)), mode="edge").astype(np.float64)
    s = b.cumsum(0).cumsum(1)
    s = np.pad(s, ((1, 0), (1, 0)))
    ny, nx = a.shape
    out = s[k:k + ny, k:k + nx] - s[0:ny, k:k + nx] - s[k:k + ny, 0:nx] + s[0:ny, 0:nx]
    return out / (k * k)


def diamond_square(k, hurst, rng, dtype=np.float32):
    """Fractal surface on a (2^k+1)^2 grid; displacement scale 2^(-hurst) per level."""
    n = (1 << k) + 1
    z = np.zeros((n, n), dtype=dtype)
    z[0, 0], z[0, -1], z[-1, 0], z[-1, -1] = rng.standard_normal(4).astype(dtype)
    step, scale = n - 1, 1.0
    while step > 1:
        h = step // 2
        c = (z[0:-1:step, 0:-1:step] + z[0:-1:step, step::step] + z[step::step, 0:-1:step]
             + z[step::step, step::step]) * dtype(0.25)
        z[h::step, h::step] = c + dtype(scale) * rng.standard_normal(c.shape, dtype=dtype)
        # square step, rows h::step / cols 0::step (vertical edge midpoints)
        for r0, c0 in ((h, 0), (0, h)):
            rows = np.arange(r0, n, step)
            cols = np.arange(c0, n, step)
            acc = np.zeros((len(rows), len(cols)), dtype=dtype)
            cnt = np.zeros((len(rows), len(cols)), dtype=dtype)
            for dr, dc in ((-h, 0), (h, 0), (0, -h), (0, h)):
                rr, cc = rows + dr, cols + dc
                okr = (rr >= 0) & (rr < n)
                okc = (cc >= 0) & (cc < n)
                sub = z[np.ix_(rr[okr], cc[okc])]
                acc[np.ix_(okr, okc)] += sub
                cnt[np.ix_(okr, okc)] += 1
            z[np.ix_(rows, cols)] = acc / cnt + dtype(scale) * rng.standard_normal(
                acc.shape, dtype=dtype)
        step = h
        scale *= 2.0 ** (-hurst)
    return z


def fractal_dem(nx, ny, relief, hurst, seed):
    rng = np.random.default_rng(seed)
    k = max(1, math.ceil(math.log2(max(nx, ny) - 1 if max(nx, ny) > 2 else 2)))
    while (1 << k) + 1 < max(nx, ny):
        k += 1
    z = diamond_square(k, hurst, rng)[:ny, :nx].astype(np.float64)
    lo, hi = float(z.min()), float(z.max())
    return (z - lo) / (hi - lo if hi > lo else 1.0) * relief


def _distinct_cells(rng, valid, n):
    """n distinct (x, y) of cells where `valid` is True."""
    ny, nx = valid.shape
    out = []
    seen = set()
    while len(out) < n:
        m = 2 * (n - len(out)) + 16
        xs = rng.integers(0, nx, m)
        ys = rng.integers(0, ny, m)
        for x, y in zip(xs.tolist(), ys.tolist()):
            if valid[y, x] and (x, y) not in seen:
                seen.add((x, y))
                out.append((x, y))
                if len(out) == n:
                    break
    return out


def plane_case(nx=256, ny=256, dx=10.0, sim_time=3600.0, seed=1):
    """Eastward plane of slope 0.001 plus N(0, 0.5) noise smoothed over 5 x 5
    cells and rescaled to sigma 0.5 m; one 5 m^3/s QFIX source at the centre;
    FREE east edge, other edges closed; n=0.035, cfl=0.7, depth_thresh=0.001."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dx
    plane = 0.001 * (x[-1] - x)[None, :].repeat(ny, 0)
    noise = box_smooth(rng.standard_normal((ny, nx)), 5)
    noise = noise / noise.std() * 0.5
    dem = plane + noise
    src = PointBC(nx // 2, ny // 2, AF_BC_QFIX, 5.0 / dx)
    return Scenario(dem=dem, dx=dx, sim_time=sim_time, east_free=True, points=[src],
                    name=f"plane {nx}x{ny}")


def valley_case(n=4096, dx=5.0, sim_time=7200.0, n_qvar=2000, seed=2):
    """Fractal DEM (hurst 0.8, 50 m relief) with a sinuous west-east valley;
    2000 QVAR triangular hydrographs (peak U(0.1, 2) m^3/s at U(600, 3000) s);
    20 mm/h rain over 0-1800 s; HFIX 2.0 m on every west-edge cell; FREE N/E/S."""
    rng = np.random.default_rng(seed)
    dem = fractal_dem(n, n, 50.0, 0.8, seed)
    x = np.arange(n)
    yc = n / 2 + (n / 8) * np.sin(2 * np.pi * x / (n / 2))
    yy = np.arange(n)[:, None]
    d = (yy - yc[None, :]) / 40.0
    dem = dem - 15.0 * np.exp(-d * d)
    dem -= dem.min()
    valid = np.ones((n, n), dtype=bool)
    pts = []
    for (cx, cy) in _distinct_cells(rng, valid, n_qvar):
        peak = rng.uniform(0.1, 2.0) / dx
        tp = rng.uniform(600.0, 3000.0)
        pts.append(PointBC(cx, cy, AF_BC_QVAR, 0.0, np.array([0.0, tp, 2 * tp]),
                           np.array([0.0, peak, 0.0])))
    for j in range(n):
        pts.append(PointBC(0, j, AF_BC_HFIX, 2.0))
    rate = 0.020 / 3600.0
    return Scenario(dem=dem, dx=dx, sim_time=sim_time, north_free=True, east_free=True,
                    south_free=True, points=pts, rain_times=np.array([0.0, 1800.0, 1800.0]),
                    rain_rates=np.array([rate, rate, 0.0]), name=f"valley {n}x{n}")


def roofline_case(n=16384, dx=2.0, sim_time=3600.0, n_qfix=5000, nodata_frac=0.10, seed=3):
    """Fractal DEM (80 m relief), NODATA blobs covering ~10% of cells
    (smoothed noise above its 90th percentile), 5000 QFIX sources of
    U(0.05, 1) m^3/s at distinct valid cells, all four edges FREE."""
    rng = np.random.default_rng(seed)
    dem = fractal_dem(n, n, 80.0, 0.8, seed)
    coarse = max(8, n // 64)
    blob = box_smooth(rng.standard_normal((coarse, coarse)), 3)
    # bilinear upsampling of the coarse field to n x n
    u = (np.arange(n) + 0.5) * coarse / n - 0.5
    i0 = np.clip(np.floor(u).astype(int), 0, coarse - 2)
    w = np.clip(u - i0, 0.0, 1.0)
    rows = blob[i0] * (1 - w)[:, None] + blob[i0 + 1] * w[:, None]
    field = rows[:, i0] * (1 - w)[None, :] + rows[:, i0 + 1] * w[None, :]
    thr = np.quantile(field[:: max(1, n // 1024), :: max(1, n // 1024)], 1.0 - nodata_frac)
    mask = field > thr
    dem[mask] = -9999.0
    pts = [PointBC(x, y, AF_BC_QFIX, rng.uniform(0.05, 1.0) / dx)
           for (x, y) in _distinct_cells(rng, ~mask, n_qfix)]
    return Scenario(dem=dem, dx=dx, sim_time=sim_time, west_free=True, east_free=True,
                    north_free=True, south_free=True, points=pts, name=f"roofline {n}x{n}")


def ensemble_member(k, seed_base=4000, dx=5.0, sim_time=3600.0, lo=128, hi=2048):
    """Scenario k of the ragged ensemble: long side log-uniform in [lo, hi],
    aspect U(1, 4), Manning n U(0.02, 0.08), rain U(0, 50) mm/h, 10-500 QVAR."""
    rng = np.random.default_rng(seed_base + k)
    L = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    a = rng.uniform(1.0, 4.0)
    S = max(16, int(round(L / a)))
    if rng.random() < 0.5:
        nx, ny = L, S
    else:
        nx, ny = S, L
    dem = fractal_dem(nx, ny, rng.uniform(10.0, 60.0), 0.8, seed_base + k)
    n_q = int(rng.integers(10, 501))
    valid = np.ones((ny, nx), dtype=bool)
    pts = []
    for (cx, cy) in _distinct_cells(rng, valid, n_q):
        peak = rng.uniform(0.1, 2.0) / dx
        tp = rng.uniform(300.0, 0.8 * sim_time)
        pts.append(PointBC(cx, cy, AF_BC_QVAR, 0.0, np.array([0.0, tp, 2 * tp]),
                           np.array([0.0, peak, 0.0])))
    rate = rng.uniform(0.0, 0.050) / 3600.0
    edges = rng.random(4) < 0.5
    return Scenario(dem=dem, dx=dx, fpn=float(rng.uniform(0.02, 0.08)), sim_time=sim_time,
                    west_free=bool(edges[0]), east_free=bool(edges[1]),
                    north_free=bool(edges[2]), south_free=bool(edges[3]), points=pts,
                    rain_times=np.array([0.0, sim_time]), rain_rates=np.array([rate, rate]),
                    name=f"ensemble {k} {nx}x{ny}")


def ensemble_shape(k, seed_base=4000, lo=128, hi=2048):
    """(nx, ny) of ensemble_member(k) from its first draws alone, without
    building the terrain (the same generator, stopped after the shape)."""
    rng = np.random.default_rng(seed_base + k)
    L = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    a = rng.uniform(1.0, 4.0)
    S = max(16, int(round(L / a)))
    return (L, S) if rng.random() < 0.5 else (S, L)


def smallest_members(count, n=512, seed_base=4000):
    """Indices of the `count` ensemble members with the fewest cells."""
    cells = [(ensemble_shape(k, seed_base)[0] * ensemble_shape(k, seed_base)[1], k)
             for k in range(n)]
    return [k for _, k in sorted(cells)[:count]]


def ensemble(n=512, seed_base=4000, sim_time=3600.0, lo=128, hi=2048):
    return [ensemble_member(k, seed_base, sim_time=sim_time, lo=lo, hi=hi) for k in range(n)]


def arrivals(n, rate=20.0, seed=5):
    """Poisson arrival times (s) and per-request sim_time U(600, 3600) s."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    sims = rng.uniform(600.0, 3600.0, n)
    return t, sims


def small_case(nx, ny, seed, dx=10.0, nodata_frac=0.05, n_src=6, n_over=4, rain=True,
               edges=(True, True, True, True), sim_time=300.0, dup=True):
    """A small random scenario exercising every path of the step: NODATA cells,
    QFIX and QVAR sources (including two on one cell), FREE/HFIX/HVAR points,
    rain, and the given FREE edges."""
    rng = np.random.default_rng(seed)
    dem = box_smooth(rng.standard_normal((ny, nx)), 3) * 2.0 + np.linspace(
        3.0, 0.0, nx)[None, :]
    mask = rng.random((ny, nx)) < nodata_frac
    dem[mask] = -9999.0
    valid = ~mask
    pts = []
    cells = _distinct_cells(rng, valid, n_src + n_over)
    for k, (x, y) in enumerate(cells[:n_src]):
        if k % 2 == 0:
            pts.append(PointBC(x, y, AF_BC_QFIX, float(rng.uniform(0.05, 0.5))))
        else:
            tp = float(rng.uniform(20.0, sim_time))
            pts.append(PointBC(x, y, AF_BC_QVAR, 0.0, np.array([0.0, tp, 2 * tp]),
                               np.array([0.0, float(rng.uniform(0.1, 0.6)), 0.0])))
    if dup and n_src >= 2:
        x, y = cells[0]
        pts.append(PointBC(x, y, AF_BC_QFIX, -0.05))  # second source on one cell
    from ._lib import AF_BC_FREE, AF_BC_HVAR
    for k, (x, y) in enumerate(cells[n_src:]):
        kind = k % 3
        if kind == 0:
            pts.append(PointBC(x, y, AF_BC_FREE))
        elif kind == 1:
            pts.append(PointBC(x, y, AF_BC_HFIX, float(rng.uniform(0.1, 0.5))))
        else:
            pts.append(PointBC(x, y, AF_BC_HVAR, 0.0, np.array([0.0, sim_time]),
                               np.array([0.05, float(rng.uniform(0.2, 0.8))])))
    rt = rr = None
    if rain:
        rt = np.array([0.0, sim_time / 2, sim_time])
        rr = np.array([1e-5, 3e-5, 0.0])
    w, e, n_, s = edges
    return Scenario(dem=dem, dx=dx, sim_time=sim_time, west_free=w, east_free=e,
                    north_free=n_, south_free=s, points=pts, rain_times=rt, rain_rates=rr,
                    name=f"small {nx}x{ny} seed {seed}")
