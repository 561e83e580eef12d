"""Continuous batching (GPU): each served result equals the scenario run alone."""
import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd.domain import Domain, make_tuning
from adaptflood_amd.serve import Pool, Request, serve

pytestmark = pytest.mark.gpu


def _scenarios():
    sizes = [(60, 41), (130, 70), (17, 90), (257, 33), (45, 45), (90, 120)]
    return [generators.small_case(nx, ny, seed=70 + k, sim_time=60.0 + 25 * k)
            for k, (nx, ny) in enumerate(sizes)]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_served_equals_alone(gpu, prec):
    scns = _scenarios()
    tun = make_tuning(tile_rows=16)
    # arrivals staggered over pool steps; the budget admits only some at once
    reqs = [Request(k, s, arrival=7 * k) for k, s in enumerate(scns)]
    pool = Pool(budget_bytes=3 << 20, precision=prec, max_scenarios=4, max_tiles=512, tuning=tun)
    done, summary = serve(reqs, pool, chunk_steps=6, clock="steps", keep_results=True)
    assert summary["requests"] == len(scns)
    for r in done:
        d = Domain(r.scenario, precision=prec, tuning=tun)
        d.run(check_every=24, graph=True)
        assert r.steps == d.info()["steps"], r.rid
        assert np.array_equal(r.result, d.read("H")), r.rid


def test_pool_capacity_and_release(gpu):
    scns = _scenarios()[:3]
    pool = Pool(budget_bytes=1 << 20, precision="fp32", max_scenarios=2, max_tiles=64,
                tuning=make_tuning(tile_rows=16))
    b0 = [pool.footprint(s)[0] for s in scns]
    assert all(b > 0 for b in b0)
    s0 = pool.admit(scns[0])
    s1 = pool.admit(scns[1])
    assert s0 is not None and s1 is not None
    assert pool.admit(scns[2]) is None          # two slots only
    st = pool.stats()
    assert st["active"] == 2 and st["tiles"] > 0
    pool.release(s0)
    assert pool.stats()["active"] == 1
    s2 = pool.admit(scns[2])
    assert s2 == s0                              # the freed slot is reused
    pool.advance(12, graph=False)
    fin, act, _ = pool.poll()
    assert act == 2


def test_served_fp64_equals_oracle_default_tiling(gpu):
    """Served in a mix with the default, per-scenario tile heights (widths a
    whole number of strips and not), every FP64 result is bitwise the
    oracle's."""
    from oracle import oracle
    sizes = [(256, 40), (100, 64), (512, 24), (33, 130)]
    scns = [generators.small_case(nx, ny, seed=90 + k, sim_time=40.0 + 20 * k)
            for k, (nx, ny) in enumerate(sizes)]
    reqs = [Request(k, s, arrival=5 * k) for k, s in enumerate(scns)]
    pool = Pool(budget_bytes=4 << 20, precision="fp64", max_scenarios=2, max_tiles=4096)
    done, summary = serve(reqs, pool, chunk_steps=6, clock="steps", keep_results=True)
    assert summary["requests"] == len(scns)
    for r in done:
        ref = oracle.run(r.scenario)
        print(f"request {r.rid}: {r.scenario.nx}x{r.scenario.ny} steps={r.steps}", flush=True)
        assert r.steps == ref["steps"], r.rid
        np.testing.assert_array_equal(r.result, ref["H"])


def test_served_fp32_strip_widths_equals_alone(gpu):
    """FP32 served with the default tile heights: widths a whole number of
    256-column strips, one column more, and narrow ones, admitted while others
    are mid-run (two slots); each result and step count is bitwise the
    scenario's alone, and within the FP32 bar of the oracle."""
    from oracle import oracle
    sizes = [(256, 40), (512, 33), (257, 20), (100, 64), (768, 9)]
    scns = [generators.small_case(nx, ny, seed=110 + k, sim_time=50.0 + 15 * k)
            for k, (nx, ny) in enumerate(sizes)]
    reqs = [Request(k, s, arrival=4 * k) for k, s in enumerate(scns)]
    pool = Pool(budget_bytes=4 << 20, precision="fp32", max_scenarios=2, max_tiles=4096)
    done, summary = serve(reqs, pool, chunk_steps=5, clock="steps", keep_results=True)
    assert summary["requests"] == len(scns)
    for r in done:
        d = Domain(r.scenario, precision="fp32")
        d.run(check_every=20, graph=True)
        ref = oracle.run(r.scenario)
        err = float(np.max(np.abs(r.result - ref["H"])))
        print(f"request {r.rid}: {r.scenario.nx}x{r.scenario.ny} steps={r.steps} "
              f"max|dH| vs oracle={err:.2e}", flush=True)
        assert r.steps == d.info()["steps"], r.rid
        assert np.array_equal(r.result, d.read("H")), r.rid
        assert err <= 1e-3


def test_request_over_budget_is_rejected(gpu):
    """With the real pool: a scenario whose arena extent exceeds the whole HBM
    budget is rejected; the others are served and equal their solo runs."""
    small = [generators.small_case(40, 30, seed=120 + k, sim_time=30.0) for k in range(2)]
    big = generators.small_case(700, 600, seed=130, sim_time=30.0)
    pool = Pool(budget_bytes=2 << 20, precision="fp32", max_scenarios=4, max_tiles=4096)
    assert pool.footprint(big)[0] > (2 << 20) > pool.footprint(small[0])[0]
    reqs = [Request(0, small[0], arrival=0), Request(1, big, arrival=1),
            Request(2, small[1], arrival=2)]
    done, summary = serve(reqs, pool, chunk_steps=6, clock="steps", keep_results=True)
    assert summary["rejected"] == 1 and reqs[1].error
    assert sorted(r.rid for r in done) == [0, 2]
    for r in done:
        d = Domain(r.scenario, precision="fp32")
        d.run(check_every=12, graph=True)
        assert np.array_equal(r.result, d.read("H")), r.rid


def test_default_tile_height_follows_pool_load(gpu):
    """Default tiling in a pool: a small scenario admitted to an empty pool
    takes the shortest tiles (2 rows: it runs alone and is latency-bound);
    admitted beside a domain that already offers 2048 tiles, it stops at 8
    rows.  The footprint query reports the larger, 2-row tile count."""
    from adaptflood_amd.scenario import Scenario
    small = generators.small_case(128, 32, seed=140, sim_time=30.0)
    dem = np.tile(np.linspace(2.0, 0.0, 2048), (2048, 1))
    big = Scenario(dem=dem, dx=5.0, sim_time=10.0)
    alone = Pool(budget_bytes=64 << 20, precision="fp32", max_scenarios=2, max_tiles=8192)
    assert alone.footprint(small)[1] == 16             # 1 strip x 32/2 segments
    alone.admit(small)
    assert alone.stats()["tiles"] == 16
    shared = Pool(budget_bytes=1 << 30, precision="fp32", max_scenarios=2, max_tiles=8192)
    shared.admit(big)
    assert shared.stats()["tiles"] == 2048             # 8 strips x 2048/8 segments
    shared.admit(small)
    assert shared.stats()["tiles"] == 2048 + 4         # 1 strip x 32/8 segments


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_pool_retiles_when_it_drains(gpu, prec):
    """A scenario admitted beside a domain that fills the launch takes 8-row
    tiles.  When that domain is released the pool holds fewer than 2048 tiles,
    and the scenario moves, mid-run and with a mass report pending, to the
    2-row tiles it would take alone.  Its depth grid and step count equal its
    alone run bitwise; its reports have the same times, steps and wet areas,
    and volumes equal to rounding (the partial sums are grouped by tile)."""
    import dataclasses
    from adaptflood_amd.domain import Batch
    from adaptflood_amd.scenario import Scenario
    small = dataclasses.replace(generators.small_case(128, 32, seed=141, sim_time=300.0),
                                mass_int=1.0)
    dem = np.tile(np.linspace(2.0, 0.0, 2048), (2048, 1))
    big = Scenario(dem=dem, dx=5.0, sim_time=1e6)
    pool = Pool(budget_bytes=2 << 30, precision=prec, max_scenarios=2, max_tiles=8192)
    sb = pool.admit(big)
    ss = pool.admit(small)
    assert pool.stats()["tiles"] == 2048 + 4            # the small one at 8 rows
    pool.advance(13, graph=False)
    assert pool.info(ss)["t"] < small.sim_time
    pool.release(sb)
    assert pool.stats()["tiles"] == 16                  # now at 2 rows
    for _ in range(10000):
        pool.advance(6, graph=True)
        fin, _, _ = pool.poll()
        if ss in fin:
            break
    assert ss in fin
    d = Domain(small, precision=prec)
    d.run(check_every=6, graph=False)
    print(f"{prec}: steps {pool.info(ss)['steps']}", flush=True)
    assert pool.info(ss)["steps"] == d.info()["steps"]
    assert np.array_equal(pool.read(ss, "H"), d.read("H"))
    rp, rd = Batch.reports(pool, ss), d.reports()
    assert rp.shape == rd.shape and len(rp) > 13
    assert np.array_equal(rp[:, [0, 1, 3]], rd[:, [0, 1, 3]])
    np.testing.assert_allclose(rp[:, 2], rd[:, 2], rtol=1e-12, atol=0)


def _poisoned(kind, seed, dx=10.0):
    """A small scenario with finite inputs that goes bad on the device:
    'inf'   a QFIX of 1e308 m^2/s (v*dx overflows: an infinite depth),
    'nan'   that source and its negative on one cell (inf - inf = NaN, which
            the H maximum never sees),
    'stall' a QFIX of 1e300 (a finite 1e300 m depth, whose dt of ~1e-150 s
            cannot move the FP64 clock past t = 10 s)."""
    import dataclasses
    from adaptflood_amd._lib import AF_BC_QFIX
    from adaptflood_amd.scenario import PointBC
    s = generators.small_case(48, 40, seed=seed, dx=dx, sim_time=200.0)
    x, y = 20, 17
    pts = list(s.points)
    if kind == "inf":
        pts.append(PointBC(x, y, AF_BC_QFIX, 1e308))
    elif kind == "nan":
        pts += [PointBC(x, y, AF_BC_QFIX, 1e308), PointBC(x, y, AF_BC_QFIX, -1e308)]
    else:
        pts.append(PointBC(x, y, AF_BC_QFIX, 1e300))
    s.dem[y, x] = 1.0
    return dataclasses.replace(s, points=pts, name=f"poisoned {kind}")


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_bad_scenarios_are_isolated_in_the_pool(gpu, prec):
    """Served beside healthy scenarios: a scenario whose depth turns infinite,
    one whose depth turns NaN, (FP64) one whose clock stalls, and one refused
    at admission for an infinite initial depth each come back with an error;
    the loop ends, and each healthy result equals its solo run bitwise."""
    import dataclasses
    from adaptflood_amd._lib import AF_SCN_NONFINITE, AF_SCN_STALLED, AFError
    healthy = [generators.small_case(nx, ny, seed=150 + k, sim_time=80.0 + 20 * k)
               for k, (nx, ny) in enumerate([(60, 41), (130, 70), (90, 33)])]
    bad = [_poisoned("inf", 160), _poisoned("nan", 161)]
    if prec == "fp64":
        bad.append(_poisoned("stall", 162))
    h0 = np.zeros_like(healthy[0].dem)
    h0[3, 4] = np.inf
    refused = dataclasses.replace(healthy[0], h0=h0, name="infinite h0")
    order = [healthy[0], bad[0], healthy[1], refused, bad[1], healthy[2]] + bad[2:]
    reqs = [Request(k, s, arrival=3 * k) for k, s in enumerate(order)]
    pool = Pool(budget_bytes=8 << 20, precision=prec, max_scenarios=4, max_tiles=4096)
    done, summary = serve(reqs, pool, chunk_steps=6, clock="steps", keep_results=True)
    for r in reqs:
        print(f"request {r.rid} ({r.scenario.name}): steps={r.steps} error={r.error}",
              flush=True)
    assert sorted(r.scenario.name for r in done) == sorted(s.name for s in healthy)
    assert summary["rejected"] == 1 and summary["failed"] == len(bad)
    by_name = {r.scenario.name: r for r in reqs}
    assert "not finite" in by_name["infinite h0"].error
    assert "NaN or infinite" in by_name["poisoned inf"].error
    assert "NaN or infinite" in by_name["poisoned nan"].error
    if prec == "fp64":
        assert "stalled" in by_name["poisoned stall"].error
    for r in done:
        d = Domain(r.scenario, precision=prec)
        d.run(check_every=12, graph=True)
        assert r.steps == d.info()["steps"], r.rid
        assert np.array_equal(r.result, d.read("H")), r.rid
    # a lone domain reports the stop as an error, with its status
    d = Domain(bad[1], precision=prec)
    with pytest.raises(AFError, match="NaN or infinite"):
        d.run(check_every=6, graph=False)
    inf = d.info()
    assert inf["status"] == AF_SCN_NONFINITE and inf["finished"] and inf["steps"] <= 2
    if prec == "fp64":
        d = Domain(bad[2], precision=prec)
        with pytest.raises(AFError, match="stalled"):
            d.run(check_every=6, graph=True)
        assert d.info()["status"] == AF_SCN_STALLED and d.info()["steps"] == 1


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_lane_steps_equal_alone(gpu, prec):
    """The pool's lane (af_pool_set_lane / af_pool_advance_lane): the lane's
    scenarios step on the lane's own stream, before, beside or after the
    pool's launches, the lane's members change mid-run, and every scenario's
    step count, depth grid, discharges and mass reports still equal its lone
    run bitwise.  A lane burst that is not a multiple of 6 steps, or a lane
    change where the pool's step index is not a multiple of 6, is refused."""
    import dataclasses
    from adaptflood_amd._lib import AFError
    from adaptflood_amd.domain import Batch
    scns = [generators.small_case(96, 40, seed=160, dx=2.0, sim_time=60.0),
            generators.small_case(64, 64, seed=161, sim_time=120.0),
            generators.small_case(700, 200, seed=162, dx=3.0, sim_time=90.0),
            generators.small_case(40, 90, seed=163, sim_time=150.0)]
    scns = [dataclasses.replace(s, mass_int=5.0) for s in scns]
    tun = make_tuning(tile_rows=16)
    pool = Pool(budget_bytes=64 << 20, precision=prec, max_scenarios=4, max_tiles=512,
                tuning=tun)
    slots = [pool.admit(s) for s in scns]
    assert None not in slots
    with pytest.raises(AFError):
        pool.advance_lane(8)
    pool.advance(3, graph=False)
    with pytest.raises(AFError):
        pool.set_lane([slots[0]])              # the pool stands at step 3
    pool.advance(3, graph=False)
    pool.set_lane([slots[0]])
    left = set(slots)
    steps = {}
    for it in range(20000):
        if it == 40:
            pool.set_lane([slots[0], slots[2], slots[3]])
        if it == 90:
            pool.set_lane([slots[2]])
        if it == 150:
            pool.set_lane([slots[1], slots[3]])
        if it % 2:
            pool.advance_lane(18, graph=(it % 3 != 0))     # beside the chunk
            pool.advance(12, graph=(it % 4 == 1))
        else:
            pool.advance(12, graph=(it % 4 == 0))
            pool.advance_lane(18, graph=(it % 3 != 0))     # behind it
        fin, _, _ = pool.poll()
        if it == 1:
            assert pool.lane_ms() > 0.0
        for s in fin:
            if s in left:
                steps[s] = pool.info(s)["steps"]
                left.discard(s)
        if not left:
            break
    assert not left
    for s, scn in zip(slots, scns):
        d = Domain(scn, precision=prec, tuning=tun)
        d.run(check_every=24, graph=True)
        print(f"{prec} {scn.name}: {steps[s]} steps", flush=True)
        assert steps[s] == d.info()["steps"]
        assert np.array_equal(pool.read(s, "H"), d.read("H")), scn.name
        assert np.array_equal(pool.read(s, "Qx"), d.read("Qx")), scn.name
        assert np.array_equal(pool.read(s, "Qy"), d.read("Qy")), scn.name
        rp, rd = Batch.reports(pool, s), d.reports()
        assert len(rd) >= 10 and np.array_equal(rp, rd), scn.name
        d.close()


def test_served_with_lane_equals_alone(gpu):
    """serve() puts a request whose time step is a fifth of the others' in the
    lane; every result still equals its lone run bitwise."""
    scns = [generators.small_case(64, 48, seed=170 + k, sim_time=120.0) for k in range(9)]
    scns.append(generators.small_case(64, 48, seed=179, dx=2.0, sim_time=120.0))
    tun = make_tuning(tile_rows=16)

    class Watched(Pool):
        lanes = []

        def set_lane(self, slots):
            self.lanes.append(tuple(slots))
            super().set_lane(slots)

    reqs = [Request(k, s, arrival=0) for k, s in enumerate(scns)]
    pool = Watched(budget_bytes=16 << 20, precision="fp32", max_scenarios=10, max_tiles=1024,
                   tuning=tun)
    done, summary = serve(reqs, pool, chunk_steps=12, clock="steps", keep_results=True,
                          lane=True)
    assert summary["requests"] == len(scns) and summary["failed"] == 0
    assert any(len(l) for l in Watched.lanes)
    for r in done:
        d = Domain(r.scenario, precision="fp32", tuning=tun)
        d.run(check_every=24, graph=True)
        assert r.steps == d.info()["steps"], r.rid
        assert np.array_equal(r.result, d.read("H")), r.rid
        d.close()


def test_staged_admission_equals_alone(gpu):
    """Scenarios staged on worker threads (af_pool_stage) and admitted with
    af_pool_admit_staged step exactly as when admitted directly; a scenario
    staged for the other precision is refused."""
    import concurrent.futures
    from adaptflood_amd.serve import RequestError, Staged
    from adaptflood_amd._lib import AF_FP64
    scns = [generators.small_case(90, 40 + 9 * k, seed=190 + k, sim_time=60.0) for k in range(4)]
    tun = make_tuning(tile_rows=16)
    pool = Pool(budget_bytes=8 << 20, precision="fp32", max_scenarios=4, max_tiles=512, tuning=tun)
    with concurrent.futures.ThreadPoolExecutor(4) as ex:
        staged = list(ex.map(pool.stage, scns))
    with pytest.raises(RequestError, match="does not match"):
        pool.admit(Staged(scns[0], AF_FP64))
    slots = [pool.admit(st) for st in staged]
    assert None not in slots and all(st.handle is None for st in staged)
    left = set(slots)
    for _ in range(5000):
        pool.advance(12, graph=True)
        fin, _, _ = pool.poll()
        left -= set(fin)
        if not left:
            break
    assert not left
    for s, scn in zip(slots, scns):
        d = Domain(scn, precision="fp32", tuning=tun)
        d.run(check_every=24, graph=True)
        assert pool.info(s)["steps"] == d.info()["steps"]
        assert np.array_equal(pool.read(s, "H"), d.read("H")), scn.name
        d.close()
