"""Fused HIP step vs the FP64 oracle (GPU).

Bars: FP64 fused path max|dH| <= 1e-9 m against the oracle; FP32 fused path
max|dH| <= 1e-3 m and relative volume error <= 1e-5.  Step counts and dt
sequences must match the oracle's exactly at FP64.  The FP64 path evaluates
the oracle's operations in the oracle's order (pow included), so its states
are also checked to be bitwise equal.
"""
import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd.domain import Batch, Domain, make_tuning
from oracle import oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
FP32_TOL = 1e-3


def _small(seed, nx, ny, **kw):
    return generators.small_case(nx, ny, seed=seed, **kw)


def _faces_close(a, b, tol):
    return float(np.max(np.abs(a - b))) <= tol


@pytest.mark.parametrize("seed,nx,ny,rows", [
    (1, 37, 29, 8), (2, 64, 64, 16), (3, 130, 70, 32), (4, 257, 19, 5), (5, 300, 300, 128),
    (6, 1100, 40, 16), (7, 256, 48, 8), (8, 512, 20, 7),
    # extreme aspects: 128 strips of three rows, one column pair down 4096
    # rows, and one-row tiles down a three-column channel
    (9, 16384, 3, 2), (10, 2, 4096, 64), (11, 3, 3000, 1),
])
def test_fp64_matches_oracle_small(gpu, seed, nx, ny, rows):
    scn = _small(seed, nx, ny, sim_time=200.0)
    print(f"case seed={seed} {nx}x{ny} rows={rows}", flush=True)
    ref = oracle.run(scn)
    d = Domain(scn, precision="fp64", tuning=make_tuning(tile_rows=rows, dt_log_cap=1 << 16))
    d.run(check_every=50, graph=False)
    info = d.info()
    assert info["steps"] == ref["steps"]
    assert info["t"] == pytest.approx(ref["t"], abs=1e-9)
    np.testing.assert_allclose(d.dt_log(), ref["dt"], rtol=0, atol=1e-12)
    assert float(np.max(np.abs(d.read("H") - ref["H"]))) <= FP64_TOL
    assert _faces_close(d.read("Qx"), ref["Qx"], 1e-8)
    assert _faces_close(d.read("Qy"), ref["Qy"], 1e-8)
    assert _faces_close(d.read("Qxold"), ref["Qxold"], 1e-9)
    np.testing.assert_array_equal(d.read("H"), ref["H"])
    np.testing.assert_array_equal(d.read("Qx"), ref["Qx"])
    np.testing.assert_array_equal(d.read("Qyold"), ref["Qyold"])
    np.testing.assert_array_equal(d.dt_log(), ref["dt"])


@pytest.mark.parametrize("edges", [(False, False, False, False), (True, False, True, False),
                                   (False, True, False, True), (True, True, True, True)])
def test_fp64_edges(gpu, edges):
    scn = _small(11, 45, 38, edges=edges, sim_time=150.0)
    print(f"edges={edges}", flush=True)
    ref = oracle.run(scn)
    d = Domain(scn, precision="fp64", tuning=make_tuning(tile_rows=7))
    d.run(check_every=40, graph=False)
    assert d.info()["steps"] == ref["steps"]
    assert float(np.max(np.abs(d.read("H") - ref["H"]))) <= FP64_TOL
    assert _faces_close(d.read("Qx"), ref["Qx"], 1e-8)
    assert _faces_close(d.read("Qy"), ref["Qy"], 1e-8)
    np.testing.assert_array_equal(d.read("H"), ref["H"])
    np.testing.assert_array_equal(d.read("Qy"), ref["Qy"])


def test_fp64_bitwise_on_unstable_flood(gpu):
    """A flood that the scheme amplifies (HFIX inflow over fractal terrain,
    depths growing to tens of metres): a 1-ulp difference anywhere grows
    about tenfold every 50 steps here, so only bitwise agreement keeps the
    FP64 path on the oracle's trajectory."""
    scn = generators.valley_case(n=384, n_qvar=20)
    scn.sim_time = 150.0
    ref = oracle.run(scn)
    print(f"valley 384: oracle steps={ref['steps']}", flush=True)
    d = Domain(scn, precision="fp64", tuning=make_tuning(dt_log_cap=1 << 16))
    d.run(check_every=72, graph=True)
    assert d.info()["steps"] == ref["steps"]
    np.testing.assert_array_equal(d.dt_log(), ref["dt"])
    np.testing.assert_array_equal(d.read("H"), ref["H"])
    np.testing.assert_array_equal(d.read("Qx"), ref["Qx"])


def test_plane_case_fp64_full_run(gpu):
    """256 x 256, 3600 s: the FP64 fused path against the oracle."""
    scn = generators.plane_case()
    ref = oracle.run(scn)
    d = Domain(scn, precision="fp64", tuning=make_tuning(dt_log_cap=1 << 16))
    d.run(check_every=360, graph=True)
    info = d.info()
    print(f"plane fp64 steps={info['steps']} oracle={ref['steps']}", flush=True)
    assert info["steps"] == ref["steps"]
    err = float(np.max(np.abs(d.read("H") - ref["H"])))
    print(f"plane fp64 max|dH|={err:.3e}", flush=True)
    assert err <= FP64_TOL
    rep = d.reports()
    assert rep.shape == ref["reports"].shape
    np.testing.assert_allclose(rep[:, 0], ref["reports"][:, 0], rtol=0, atol=1e-9)
    np.testing.assert_allclose(rep[:, 1], ref["reports"][:, 1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rep[:, 2], ref["reports"][:, 2], rtol=1e-12)
    np.testing.assert_allclose(rep[:, 3], ref["reports"][:, 3], rtol=0, atol=1e-9)


@pytest.mark.parametrize("nx", [256, 512, 513])
def test_fp32_strip_multiple_widths(gpu, nx):
    """Grids whose width is a whole number of FP32 strips (256 columns), and
    one column more: the east boundary face sits just outside the last strip
    (read through the strip-edge halo) or inside it."""
    scn = _small(70 + nx % 7, nx, 40, sim_time=120.0)
    ref = oracle.run(scn)
    d = Domain(scn, precision="fp32", tuning=make_tuning(tile_rows=8))
    d.run(check_every=30, graph=False)
    err = float(np.max(np.abs(d.read("H") - ref["H"])))
    print(f"fp32 {nx}x40: max|dH|={err:.2e}", flush=True)
    assert err <= FP32_TOL


def test_plane_case_fp32_full_run(gpu):
    scn = generators.plane_case()
    ref = oracle.run(scn)
    d = Domain(scn, precision="fp32")
    d.run(check_every=360, graph=True)
    H = d.read("H")
    err = float(np.max(np.abs(H - ref["H"])))
    vol, _ = d.volume()
    vref = float(np.sum(ref["H"]) * scn.dx * scn.dx)
    rel = abs(vol - vref) / vref
    print(f"plane fp32 max|dH|={err:.3e} rel vol err={rel:.3e} steps={d.info()['steps']} "
          f"(oracle {ref['steps']})", flush=True)
    assert err <= FP32_TOL
    assert rel <= 1e-5


def test_graph_equals_eager(gpu):
    scn = _small(21, 90, 77, sim_time=300.0)
    a = Domain(scn, precision="fp32", tuning=make_tuning(tile_rows=16, graph_steps=12))
    a.run(check_every=30, graph=True)
    b = Domain(scn, precision="fp32", tuning=make_tuning(tile_rows=16))
    b.run(check_every=30, graph=False)
    assert a.info()["steps"] == b.info()["steps"]
    assert np.array_equal(a.read("H"), b.read("H"))
    assert np.array_equal(a.read("Qx"), b.read("Qx"))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_tiling_invariance(gpu, prec):
    """Results do not depend on the tile shape, the kernel variant or the tile
    order (bitwise); a wide grid (nx = 2100) takes the 1 KiB row pitch."""
    scn = _small(31, 2100, 141, sim_time=60.0)
    outs = []
    for rows, wpb, var, order in ((4, 1, 1, 0), (32, 4, 0, 0), (128, 4, 2, 0), (61, 4, 3, 1),
                                  (33, 1, 2, 1), (2, 4, 2, 0), (1, 1, 1, 0)):
        d = Domain(scn, precision=prec,
                   tuning=make_tuning(tile_rows=rows, waves_per_block=wpb, variant=var,
                                      tile_order=order))
        d.run(check_every=25, graph=False)
        outs.append((d.read("H"), d.read("Qx"), d.read("Qy")))
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)


def test_batch_equals_solo(gpu):
    """Each scenario of a ragged batch, its tiles in largest-first order, gives
    bitwise the result it gives alone."""
    scns = [_small(40 + k, nx, ny, sim_time=100.0 + 30 * k)
            for k, (nx, ny) in enumerate([(50, 40), (130, 21), (17, 90), (260, 64)])]
    b = Batch(scns, precision="fp32", tuning=make_tuning(tile_rows=16, tile_order=1))
    b.run(check_every=32, graph=True)
    for k, s in enumerate(scns):
        d = Domain(s, precision="fp32", tuning=make_tuning(tile_rows=16))
        d.run(check_every=32, graph=True)
        assert b.info(k)["steps"] == d.info()["steps"]
        assert np.array_equal(b.read("H", k), d.read("H"))
        assert np.array_equal(b.read("Qy", k), d.read("Qy"))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_stop_and_resume_equals_uninterrupted(gpu, prec):
    """Moving the stop time (set_sim_time) to stop at intermediate times and
    resume gives bitwise the uninterrupted run: skipped launches leave the
    state, the clock, the H-max slots and the reports as they were."""
    scn = _small(51, 70, 45, sim_time=240.0)
    scn.mass_int = 20.0
    a = Domain(scn, precision=prec, tuning=make_tuning(tile_rows=16, graph_steps=12))
    a.run(check_every=24, graph=True)
    b = Domain(scn, precision=prec, tuning=make_tuning(tile_rows=16, graph_steps=12))
    for stop in (35.0, 36.0, 90.0, 170.0, 240.0):
        b.set_sim_time(stop)
        b.run(check_every=7, graph=True)
        assert b.info()["t"] >= stop
        b.advance(5, graph=False)  # launches past the stop time are no-ops
    ia, ib = a.info(), b.info()
    assert ia["steps"] == ib["steps"] and ia["t"] == ib["t"]
    for what in ("H", "Qx", "Qy"):
        assert np.array_equal(a.read(what), b.read(what)), what
    assert np.array_equal(a.reports(), b.reports())


def test_autotune_picks_a_candidate_and_caches(gpu, tmp_path):
    from adaptflood_amd.tune import autotune
    scns = [_small(60 + k, nx, ny, sim_time=50.0) for k, (nx, ny) in
            enumerate([(300, 40), (90, 130)])]
    cache = str(tmp_path / "tune.json")
    tun, table = autotune(scns, "fp32", rows=(8, 32), variants=(1, 2), orders=(0, 1),
                          launches=5, cache_path=cache)
    assert len(table) == 8
    assert (tun.tile_rows, tun.variant, tun.tile_order) in {(r, v, o) for _, r, v, _, o in table}
    assert table[0][1] == tun.tile_rows and table[0][2] == tun.variant
    assert table[0][4] == tun.tile_order
    again, table2 = autotune(scns, "fp32", rows=(8, 32), variants=(1, 2), cache_path=cache)
    assert (again.tile_rows, again.variant, again.tile_order) == \
        (tun.tile_rows, tun.variant, tun.tile_order)
    assert len(table2) == 8


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_read_cells_equals_full_read(gpu, prec):
    """The gauge read (read_cells) returns the same values as a full read of H
    at those cells, for both state parities, and NaN outside the grid."""
    scn = _small(61, 300, 37, sim_time=90.0)
    d = Domain(scn, precision=prec, tuning=make_tuning(tile_rows=8))
    n = scn.nx * scn.ny
    idx = np.array([0, 1, scn.nx - 1, scn.nx, n // 2 + 7, n - 1, -1, n, 5 * scn.nx + 3],
                   dtype=np.int64)
    for steps in (3, 4):   # odd and even step counts end on different buffers
        d.advance(steps, graph=False)
        got = d.read_cells(idx)
        H = d.read("H").ravel()
        inside = (idx >= 0) & (idx < n)
        assert np.array_equal(got[inside], H[idx[inside]])
        assert np.all(np.isnan(got[~inside]))
    assert d.read_cells(np.empty(0, dtype=np.int64)).size == 0


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_snapshot_equals_read_and_velocity(gpu, prec):
    """The device output snapshot (af_batch_snapshot) enqueued behind steps
    that have not run yet gives H equal to a full read, and Vx, Vy equal to
    the oracle's velocity (kernels.cpp:293-351) of that state; at both state
    parities, and without velocity."""
    scn = _small(81, 150, 47, sim_time=400.0)
    d = Domain(scn, precision=prec, tuning=make_tuning(tile_rows=8))
    nod = scn.nodata
    for steps in (30, 11):                     # 30, then 41 steps: both parities
        d.advance(steps, graph=False)          # enqueued; the snapshot follows on the stream
        H, Vx, Vy = d.snapshot(velocity=True)
        assert np.array_equal(H, d.read("H"))
        Qx, Qy = d.read("Qx"), d.read("Qy")
        vx, vy = oracle.velocity(H.ravel(), scn.dem.ravel().astype(np.float64), Qx, Qy,
                                 scn.nx, scn.ny, scn.dx, nod)
        print(f"{prec} after {d.info()['steps']} steps: moving cells "
              f"{int(np.count_nonzero(vx))}, {int(np.count_nonzero(vy))}", flush=True)
        assert np.count_nonzero(vx) > 1000 and np.count_nonzero(vy) > 1000
        assert np.array_equal(Vx.ravel(), vx) and np.array_equal(Vy.ravel(), vy)
    (H2,) = d.snapshot(velocity=False)
    assert np.array_equal(H2, H)


def _degenerate(kind):
    """Inputs at the edges of what the step accepts."""
    from adaptflood_amd.scenario import PointBC, Scenario
    from adaptflood_amd._lib import AF_BC_FREE, AF_BC_HFIX, AF_BC_QFIX, AF_BC_QVAR
    rng = np.random.default_rng(5)
    if kind == "2x2":                  # the smallest grid, a source and all edges free
        return Scenario(dem=np.array([[0.3, 0.1], [0.2, 0.0]]), dx=5.0, sim_time=200.0,
                        west_free=True, east_free=True, north_free=True, south_free=True,
                        points=[PointBC(0, 0, AF_BC_QFIX, 0.4)])
    if kind in ("2x300", "300x2"):     # one strip-edge column pair / two rows of strips
        ny, nx = (300, 2) if kind == "2x300" else (2, 300)
        dem = rng.uniform(0, 1, (ny, nx))
        return Scenario(dem=dem, dx=4.0, sim_time=120.0, east_free=True, south_free=True,
                        points=[PointBC(1, 1, AF_BC_QFIX, 0.3), PointBC(0, ny - 1, AF_BC_HFIX, 0.5)],
                        rain_times=np.array([0.0, 100.0]), rain_rates=np.array([2e-5, 0.0]))
    if kind == "all_nodata":           # no valid cell: dt stays InitTstep, sources still add
        dem = np.full((9, 70), -9999.0)
        return Scenario(dem=dem, dx=10.0, sim_time=95.0, west_free=True, north_free=True,
                        points=[PointBC(3, 4, AF_BC_QFIX, 0.2),
                                PointBC(60, 2, AF_BC_QVAR, 0.0, np.array([0.0, 50.0]),
                                        np.array([0.0, 1.0])),
                                PointBC(10, 8, AF_BC_HFIX, 0.7), PointBC(11, 8, AF_BC_FREE)],
                        rain_times=np.array([0.0]), rain_rates=np.array([1e-4]))
    if kind == "still_water":          # no points, no rain: water at rest in a basin
        dem = np.full((40, 90), 2.0)
        dem[5:35, 5:85] = rng.integers(0, 64, (30, 80)) / 64.0
        return Scenario(dem=dem, dx=2.0, sim_time=300.0, h0=np.where(dem < 2.0, 1.5 - dem, 0.0))
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["2x2", "2x300", "300x2", "all_nodata", "still_water"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_degenerate_inputs(gpu, kind, prec):
    """The smallest grid, one- and two-strip shapes, an all-NODATA grid (the
    H max is empty, so dt stays InitTstep while sources and overrides still
    act) and still water with no points and no rain: FP64 bitwise equal to the
    oracle, FP32 within its bar."""
    scn = _degenerate(kind)
    ref = oracle.run(scn)
    d = Domain(scn, precision=prec)
    d.run(check_every=16, graph=True)
    H = d.read("H")
    err = float(np.max(np.abs(H - ref["H"])))
    print(f"{kind} {prec}: steps={d.info()['steps']} (oracle {ref['steps']}) max|dH|={err:.2e}",
          flush=True)
    if kind == "all_nodata":
        assert np.all(ref["dt"] == scn.init_tstep) and ref["H"][4, 3] > 0
    if kind == "still_water":
        assert np.array_equal(ref["H"], scn.h0)
    if prec == "fp64":
        assert d.info()["steps"] == ref["steps"]
        np.testing.assert_array_equal(H, ref["H"])
        np.testing.assert_array_equal(d.read("Qx"), ref["Qx"])
        np.testing.assert_array_equal(d.read("Qy"), ref["Qy"])
    else:
        assert err <= FP32_TOL


def _dense_points_case():
    """A 70 x 64 grid with 300 sources and 200 overrides (duplicates on shared
    cells included), enough that one 64-row tile holds more than the 128
    entries per list that a tile evaluates at its start."""
    from adaptflood_amd.scenario import PointBC
    from adaptflood_amd._lib import AF_BC_FREE, AF_BC_HFIX, AF_BC_HVAR, AF_BC_QFIX, AF_BC_QVAR
    scn = _small(97, 70, 64, sim_time=120.0, n_src=0, n_over=0, rain=True)
    rng = np.random.default_rng(97)
    valid = np.argwhere(scn.dem != scn.nodata)
    pts = []
    for k in range(300):
        y, x = valid[rng.integers(len(valid))] if k % 7 else valid[k % 5]   # repeats
        if k % 3 == 0:
            pts.append(PointBC(int(x), int(y), AF_BC_QFIX, float(rng.uniform(-0.05, 0.4))))
        else:
            tp = float(rng.uniform(10.0, 100.0))
            pts.append(PointBC(int(x), int(y), AF_BC_QVAR, 0.0, np.array([0.0, tp, 2 * tp]),
                               np.array([0.0, float(rng.uniform(0.05, 0.5)), 0.0])))
    for k in range(200):
        y, x = valid[rng.integers(len(valid))] if k % 9 else valid[7 + k % 3]
        kind = k % 3
        if kind == 0:
            pts.append(PointBC(int(x), int(y), AF_BC_FREE))
        elif kind == 1:
            pts.append(PointBC(int(x), int(y), AF_BC_HFIX, float(rng.uniform(0.1, 0.6))))
        else:
            pts.append(PointBC(int(x), int(y), AF_BC_HVAR, 0.0, np.array([0.0, 120.0]),
                               np.array([0.05, float(rng.uniform(0.2, 0.9))])))
    scn.points = pts
    return scn


@pytest.mark.parametrize("prec,rows", [("fp64", 64), ("fp64", 16), ("fp32", 64)])
def test_tile_with_more_points_than_the_batch(gpu, prec, rows):
    """A tile whose point lists overflow the per-tile batch takes the rest one
    at a time: FP64 stays bitwise equal to the oracle (file order on shared
    cells included) and FP32 within its bar."""
    scn = _dense_points_case()
    ref = oracle.run(scn)
    d = Domain(scn, precision=prec, tuning=make_tuning(tile_rows=rows))
    d.run(check_every=20, graph=True)
    H = d.read("H")
    err = float(np.max(np.abs(H - ref["H"])))
    print(f"{prec} rows={rows}: steps={d.info()['steps']} (oracle {ref['steps']}) "
          f"max|dH|={err:.2e}", flush=True)
    if prec == "fp64":
        assert d.info()["steps"] == ref["steps"]
        np.testing.assert_array_equal(H, ref["H"])
        np.testing.assert_array_equal(d.read("Qx"), ref["Qx"])
    else:
        assert err <= FP32_TOL


def test_many_scenarios_poll_and_run_to_their_own_ends(gpu):
    """48 ragged scenarios, more than one block of the control-block gather:
    a batch run stops when every scenario has reached its own stop time, each
    equal (steps, H) to its solo run; a pool reports each slot finished in the
    first poll after its clock reaches its stop time, and never before."""
    from adaptflood_amd.serve import Pool
    scns = [_small(300 + k, 20 + 7 * (k % 5), 12 + 5 * (k % 4), sim_time=30.0 + 4.0 * k)
            for k in range(48)]
    b = Batch(scns, precision="fp32", tuning=make_tuning(graph_steps=12))
    b.run(check_every=12, graph=True)
    for k, s in enumerate(scns):
        inf = b.info(k)
        assert inf["finished"] and inf["t"] >= s.sim_time, k
    for k in range(0, 48, 8):
        d = Domain(scns[k], precision="fp32")
        d.run(check_every=12, graph=True)
        assert b.info(k)["steps"] == d.info()["steps"], k
        assert np.array_equal(b.read("H", k), d.read("H")), k
    pool = Pool(budget_bytes=64 << 20, precision="fp32", max_scenarios=64, max_tiles=1 << 14)
    slots = [pool.admit(s) for s in scns]
    assert None not in slots
    reported = {}
    for it in range(4000):
        pool.advance(6, graph=False)
        fin, act, _ = pool.poll()
        assert act == 48
        for sl in slots:
            done = pool.info(sl)["t"] >= scns[slots.index(sl)].sim_time
            assert (sl in fin) == done, (it, sl)
            if done and sl not in reported:
                reported[sl] = it
        if len(reported) == 48:
            break
    assert len(reported) == 48


def test_fp64_variant_follows_the_wet_area(gpu):
    """FP64 with the default tuning: a large domain (8192 tiles of 128 x 32
    cells) starts on the two-row variant while it is dry and moves to the
    one-row variant once a mass report counts more than 2% of its cells wet
    (here after the first step of heavy rain).  The same run with either
    variant fixed gives bitwise the same states, steps and reports."""
    from adaptflood_amd.scenario import Scenario
    ny, nx = 2048, 16384
    rng = np.random.default_rng(21)
    dem = np.linspace(8.0, 0.0, nx)[None, :] + rng.uniform(0.0, 0.5, (ny, nx))
    scn = Scenario(dem=dem, dx=5.0, sim_time=40.0, mass_int=5.0, east_free=True,
                   rain_times=np.array([0.0, 1e4]), rain_rates=np.array([2e-3, 2e-3]),
                   name="rain on a tilted plane, 16384 x 2048")
    d = Domain(scn, precision="fp64")
    assert d.variant() == 2
    d.run(check_every=6, graph=True)
    assert d.variant() == 1
    H, info, rep = d.read("H"), d.info(), d.reports()
    d.close()
    for v in (1, 2):
        f = Domain(scn, precision="fp64", tuning=make_tuning(variant=v))
        f.run(check_every=6, graph=True)
        assert f.variant() == v
        assert f.info()["steps"] == info["steps"] and f.info()["t"] == info["t"]
        assert np.array_equal(f.read("H"), H), v
        assert np.array_equal(f.reports(), rep), v
        f.close()
