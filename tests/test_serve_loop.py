"""The continuous-batching request loop (adaptflood_amd.serve.serve) on a
stand-in pool that simulates admission capacity and step counts on the CPU:
admission order, waiting while full, release and re-admission, completion
times and the latency summary, and the isolation of bad requests (refused at
admission, stopped on the device, or with a clock that stalls), and the lane
that lets a request needing many more steps than the others take extra steps
between chunks."""
from adaptflood_amd._lib import AF_SCN_NONFINITE
from adaptflood_amd.serve import Request, RequestError, serve


class Scn:
    def __init__(self, cells, steps, stall_at=None, fail_at=None, bad=False):
        self.cells, self.steps = cells, steps
        self.stall_at = stall_at    # the clock stops moving after this many steps
        self.fail_at = fail_at      # the device stops it (non-finite) after this many
        self.bad = bad              # refused at admission


class FakePool:
    """Holds up to `slots` scenarios; each finishes after its own number of
    steps; every call is logged."""

    def __init__(self, slots):
        self.slots, self.active, self.log, self.step = slots, {}, [], 0

    def admit(self, scn):
        if scn.bad:
            raise RequestError("initial depth at column 0, row 0 is not finite (inf)")
        if len(self.active) >= self.slots:
            return None
        slot = min(set(range(self.slots)) - set(self.active))
        self.active[slot] = [scn, 0]
        self.log.append(("admit", self.step, slot))
        return slot

    def release(self, slot):
        self.log.append(("release", self.step, slot))
        del self.active[slot]

    def advance(self, steps, graph=True):
        self.step += steps
        for rec in self.active.values():
            scn = rec[0]
            end = scn.steps if scn.fail_at is None else min(scn.steps, scn.fail_at)
            rec[1] = min(end, rec[1] + steps)

    def _clock(self, scn, k):
        # simulated time: one second per step, frozen past stall_at
        return float(k if scn.stall_at is None else min(k, scn.stall_at))

    def _status(self, scn, k):
        return AF_SCN_NONFINITE if (scn.fail_at is not None and k >= scn.fail_at) else 0

    def poll(self):
        done = [s for s, (scn, k) in sorted(self.active.items())
                if k >= scn.steps or self._status(scn, k)]
        clocks = {s: (self._clock(scn, k), k) for s, (scn, k) in self.active.items()}
        return done, len(self.active), clocks

    def info(self, slot):
        scn, k = self.active[slot]
        return {"steps": k, "t": self._clock(scn, k), "status": self._status(scn, k)}

    def read(self, slot, what="H"):
        return ("result", self.active[slot][0].steps)

    def synchronize(self):
        pass


def test_admission_waits_for_free_slots_and_keeps_arrival_order():
    # five requests, two slots; arrivals on the step clock
    steps = [10, 40, 20, 5, 30]
    reqs = [Request(k, Scn(100 * (k + 1), steps[k]), arrival=a)
            for k, a in enumerate([0, 0, 3, 4, 50])]
    pool = FakePool(slots=2)
    done, summary = serve(reqs, pool, chunk_steps=5, clock="steps", keep_results=True)
    assert summary["requests"] == 5
    admits = [e for e in pool.log if e[0] == "admit"]
    # requests 0 and 1 at step 0; 2 once request 0 is released; 3 after that
    assert [e[1] for e in admits][:2] == [0, 0]
    by_id = {r.rid: r for r in done}
    assert by_id[2].admitted >= by_id[0].done
    assert by_id[3].admitted >= by_id[2].admitted
    # every request ran exactly its own number of steps, and its result came back
    for r in done:
        assert r.steps == steps[r.rid]
        assert r.result == ("result", steps[r.rid])
        assert r.done >= r.arrival and r.admitted >= r.arrival
    assert summary["cell_updates"] == sum(100 * (k + 1) * steps[k] for k in range(5))
    lat = sorted(r.done - r.arrival for r in done)
    assert summary["p50_latency"] == lat[2]


def test_idle_pool_jumps_to_the_next_arrival():
    reqs = [Request(0, Scn(10, 5), arrival=0), Request(1, Scn(10, 5), arrival=1000)]
    pool = FakePool(slots=4)
    done, summary = serve(reqs, pool, chunk_steps=5, clock="steps")
    assert [r.rid for r in done] == [0, 1]
    assert done[1].admitted == 1000
    assert pool.step < 1100   # no stepping through the idle gap


class CellPool(FakePool):
    """FakePool that also holds at most `cells` cells at once."""

    def __init__(self, slots, cells):
        super().__init__(slots)
        self.cells = cells

    def admit(self, scn):
        used = sum(rec[0].cells for rec in self.active.values())
        if used + scn.cells > self.cells:
            return None
        return super().admit(scn)


def test_request_larger_than_the_pool_is_rejected_not_waited_for():
    """A request that even an empty pool cannot hold is rejected with an
    error; those behind it still run, and the loop ends."""
    reqs = [Request(0, Scn(500, 10), arrival=0), Request(1, Scn(5000, 10), arrival=1),
            Request(2, Scn(400, 10), arrival=2), Request(3, Scn(600, 5), arrival=30)]
    pool = CellPool(slots=4, cells=1000)
    done, summary = serve(reqs, pool, chunk_steps=5, clock="steps")
    assert summary["requests"] == 3 and summary["rejected"] == 1
    assert sorted(r.rid for r in done) == [0, 2, 3]
    assert reqs[1].error and reqs[1].admitted is None and reqs[1].done is None
    assert all(r.error is None for r in done) and all(r.steps > 0 for r in done)
    # admission is in arrival order: request 1 could not be judged too large
    # while request 0 still held part of the pool, so request 2 waited behind
    # it until request 0 finished and request 1 was rejected
    assert reqs[2].admitted >= reqs[0].done


def test_bad_requests_are_isolated_and_the_rest_complete():
    """A request refused at admission, one the device stops (non-finite
    depth) and one whose clock stalls each come back with an error; the loop
    ends, and every healthy request runs its own number of steps."""
    scns = [Scn(100, 30), Scn(100, 40, bad=True), Scn(100, 50, fail_at=12),
            Scn(100, 60, stall_at=20), Scn(100, 25)]
    reqs = [Request(k, s, arrival=k) for k, s in enumerate(scns)]
    pool = FakePool(slots=3)
    done, summary = serve(reqs, pool, chunk_steps=5, clock="steps", stall_polls=3)
    assert sorted(r.rid for r in done) == [0, 4]
    assert summary["requests"] == 2 and summary["rejected"] == 1 and summary["failed"] == 2
    assert "refused" in reqs[1].error and "not finite" in reqs[1].error
    assert reqs[1].admitted is None
    assert "NaN or infinite" in reqs[2].error and reqs[2].steps == 12
    assert "stalled" in reqs[3].error and reqs[3].steps > 20
    assert all(r.error is None and r.steps == scns[r.rid].steps for r in done)
    assert summary["cell_updates"] == 100 * (30 + 25)
    # every admitted request was released, so the pool is empty again
    assert not pool.active
    assert sum(1 for e in pool.log if e[0] == "release") == 4


class LaneScn:
    def __init__(self, cells, dt, sim_time):
        self.cells, self.dt, self.sim_time = cells, dt, sim_time
        self.steps = int(round(sim_time / dt))
        self.stall_at = self.fail_at = None
        self.bad = False


class LanePool(FakePool):
    """FakePool with a lane: set_lane(slots) moves slots out of the pool's
    steps (at a pool step that is a multiple of 6) and advance_lane(steps)
    steps only them, in multiples of 6."""

    def __init__(self, slots):
        super().__init__(slots)
        self.lane, self.lane_steps = [], 0

    def _clock(self, scn, k):
        return k * scn.dt if hasattr(scn, "dt") else super()._clock(scn, k)

    def set_lane(self, slots):
        assert all(s in self.active for s in slots) and self.step % 6 == 0
        self.lane = list(slots)
        self.log.append(("lane", self.step, tuple(slots)))

    def advance(self, steps, graph=True):
        self.step += steps
        for s, rec in self.active.items():
            if s not in self.lane:
                rec[1] = min(rec[0].steps, rec[1] + steps)

    def advance_lane(self, steps, graph=True):
        assert steps > 0 and steps % 6 == 0
        for s in self.lane:
            rec = self.active[s]
            rec[1] = min(rec[0].steps, rec[1] + steps)
        self.lane_steps += steps

    def lane_ms(self):
        return 0.0

    def release(self, slot):
        super().release(slot)
        self.lane = [s for s in self.lane if s != slot]


def test_lane_takes_the_slow_request_ahead_of_the_pool():
    """Nine requests of 600 steps and one of 6000 (a tenth of the time step):
    with the lane, the slow one takes several chunks of steps per chunk while
    the others run, so it ends after far fewer pool launches than its own
    6000 steps; every request still runs exactly its own number of steps, the
    fast ones are never in the lane and end when they did without it."""
    def run(lane):
        scns = [LaneScn(100, 1.0, 600.0) for _ in range(9)] + [LaneScn(100, 0.1, 600.0)]
        reqs = [Request(k, s, arrival=0) for k, s in enumerate(scns)]
        pool = LanePool(slots=10)
        done, summary = serve(reqs, pool, chunk_steps=36, clock="steps", lane=lane)
        assert summary["requests"] == 10 and summary["failed"] == 0
        assert all(r.steps == scns[r.rid].steps for r in done)
        return reqs, pool

    plain, pool0 = run(False)
    laned, pool1 = run(True)
    assert pool0.lane_steps == 0 and not any(e[0] == "lane" for e in pool0.log)
    assert pool1.lane_steps > 0
    assert all(e[2] in ((), (9,)) for e in pool1.log if e[0] == "lane")
    assert plain[9].done >= 6000
    assert pool1.lane_steps >= 4 * 36 * (600 // 36 - 2)   # from the second poll on
    assert laned[9].done <= 0.6 * plain[9].done
    # the fast requests end when they did without the lane
    assert [r.done for r in laned[:9]] == [r.done for r in plain[:9]]


def test_lane_holds_no_more_than_its_share_of_cells():
    """Two slow requests, each a fifth of the active cells: with the lane
    capped at a tenth of the cells neither joins; at a half both do."""
    scns = [LaneScn(100, 1.0, 600.0) for _ in range(3)] + [LaneScn(100, 0.1, 600.0)] * 2
    for share, joined in ((0.1, set()), (0.5, {3, 4})):
        reqs = [Request(k, s, arrival=0) for k, s in enumerate(scns)]
        pool = LanePool(slots=5)
        done, _ = serve(reqs, pool, chunk_steps=36, clock="steps", lane=True, lane_cells=share)
        got = set(s for e in pool.log if e[0] == "lane" for s in e[2])
        assert got == joined
        assert all(r.steps == scns[r.rid].steps for r in done)
