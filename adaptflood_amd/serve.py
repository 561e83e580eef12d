"""Continuous batching of flood scenarios on one GPU.

A Pool is one device arena (a torch uint8 tensor) managed by the native pool
(af_pool_*): scenarios are admitted while the HBM budget, the scenario slots
and the tile table have room, every admitted scenario advances by the same
fused step launches (each with its own adaptive dt and clock), and finished
scenarios are read back and released between launches so waiting requests
can take their place.
"""
import collections
import concurrent.futures
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import AF_E_CAPACITY, AF_E_INVALID, AF_SCN_NONFINITE, AF_SCN_STALLED, call
from .domain import PRECISIONS, _stream, _torch, make_tuning


class RequestError(ValueError):
    """A scenario the pool refuses for what it is (a non-finite input, a bad
    grid), as opposed to a pool that has no room for it right now."""


def status_text(status):
    if status & AF_SCN_NONFINITE:
        return "a step produced a NaN or infinite depth"
    if status & AF_SCN_STALLED:
        return "the clock stalled (t + dt == t)"
    return f"status {status}"


class Pool:
    def __init__(self, budget_bytes, precision="fp32", max_scenarios=1024, max_tiles=1 << 20,
                 tuning=None, device=None, stream=None):
        torch = _torch()
        self._torch = torch
        self.precision = PRECISIONS[precision]
        self.tuning = tuning if tuning is not None else make_tuning()
        n = C.c_size_t()
        call("af_pool_arena_bytes", int(max_scenarios), int(max_tiles), C.c_size_t(int(budget_bytes)),
             C.byref(n))
        self.arena_bytes = int(n.value)
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.arena = torch.empty(self.arena_bytes + 256, dtype=torch.uint8, device=dev)
        aligned = (self.arena.data_ptr() + 255) // 256 * 256
        self._stream = stream
        h = C.c_void_p()
        call("af_pool_create", self.precision, C.byref(self.tuning), int(max_scenarios),
             int(max_tiles), C.c_void_p(aligned), C.c_size_t(self.arena_bytes), C.byref(h))
        self.handle = h
        self.packed = {}

    def _s(self):
        return _stream(self._torch, self._stream)

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().af_domain_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, scn):
        """(device bytes, tile-table entries) the scenario needs."""
        pk = scn.pack()
        b = C.c_size_t()
        t = C.c_int32()
        call("af_pool_scenario_bytes", C.byref(pk.desc), self.precision, C.byref(self.tuning),
             C.byref(b), C.byref(t))
        return int(b.value), int(t.value)

    def stage(self, scn):
        """The host-only part of an admission (af_pool_stage): validation and
        the inputs converted to the device format.  Touches neither the pool
        nor the GPU, so it may run on another thread while the pool steps.
        Raises RequestError for a scenario the pool would refuse."""
        return Staged(scn, self.precision)

    def admit(self, scn):
        """Slot index, or None when the pool has no room right now; raises
        RequestError for a scenario the pool refuses (bad input).  `scn` is a
        Scenario or a Staged one (from stage())."""
        slot = C.c_int32()
        staged = None
        if isinstance(scn, Staged):
            staged, scn = scn, scn.scenario
            pk = staged.pk
            rc = _lib.lib().af_pool_admit_staged(self.handle, staged.handle, self._s(),
                                                 C.byref(slot))
        else:
            pk = scn.pack()
            rc = _lib.lib().af_pool_admit(self.handle, C.byref(pk.desc), self._s(),
                                          C.byref(slot))
        if rc == AF_E_CAPACITY:
            return None
        if rc == AF_E_INVALID:
            raise RequestError(_lib.lib().af_last_error().decode(errors="replace"))
        _lib.check("af_pool_admit", rc)
        self.packed[slot.value] = (scn, pk)
        if staged is not None:
            staged.free()       # the pool holds its own copy now
        return slot.value

    def release(self, slot):
        call("af_pool_release", self.handle, int(slot), self._s())
        self.packed.pop(slot, None)

    def advance(self, steps, graph=True):
        call("af_domain_advance", self.handle, int(steps), 1 if graph else 0, self._s())

    def synchronize(self):
        self._torch.cuda.synchronize()

    def poll(self):
        """(finished slots, number of active slots, {active slot: (t, steps)}),
        from one gather of the active control blocks.  A slot is finished when
        it reached its sim_time or a non-zero status stopped it."""
        cap = max(1, len(self.packed))
        slots = (C.c_int32 * cap)()
        t = (C.c_double * cap)()
        steps = (C.c_int64 * cap)()
        fin = (C.c_int32 * cap)()
        n = C.c_int32()
        call("af_pool_progress", self.handle, slots, t, steps, None, fin, cap, C.byref(n),
             self._s())
        m = min(n.value, cap)
        clocks = {slots[a]: (t[a], steps[a]) for a in range(m)}
        return [slots[a] for a in range(m) if fin[a]], n.value, clocks

    def set_lane(self, slots):
        """Puts the listed slots (and only them) in the pool's lane."""
        arr = (C.c_int32 * max(1, len(slots)))(*slots)
        call("af_pool_set_lane", self.handle, arr, len(slots), self._s())

    def advance_lane(self, steps, graph=True):
        """Enqueues `steps` steps (a multiple of 6) of the lane's scenarios on
        the pool's lane stream, beside the advance() that follows."""
        call("af_pool_advance_lane", self.handle, int(steps), 1 if graph else 0, self._s())

    def lane_ms(self):
        """Duration of the last lane work, in ms (waits for it to end)."""
        ms = C.c_float()
        call("af_pool_lane_ms", self.handle, C.byref(ms))
        return float(ms.value)

    def stats(self):
        a, t = C.c_int32(), C.c_int32()
        hf = C.c_size_t()
        k = C.c_int64()
        call("af_pool_stats", self.handle, C.byref(a), C.byref(t), C.byref(hf), C.byref(k))
        return {"active": a.value, "tiles": t.value, "heap_free": int(hf.value),
                "step_index": k.value}

    def info(self, slot):
        r = _lib.RunInfo()
        call("af_batch_info", self.handle, int(slot), C.byref(r), self._s())
        return {"t": r.t, "steps": r.steps_done, "last_dt": r.last_dt, "hmax": r.hmax,
                "status": r.status}

    def read(self, slot, what="H"):
        which = {"H": 0, "Qx": 1, "Qy": 2, "Qxold": 3, "Qyold": 4}[what]
        scn = self.packed[slot][0]
        if which == 0:
            out = np.empty((scn.ny, scn.nx))
        else:
            out = np.empty((scn.ny + 1) * (scn.nx + 1))
        call("af_batch_read", self.handle, int(slot), which, _lib.dptr(out), self._s())
        return out


class Staged:
    """A scenario staged for admission (af_pool_stage), with the packed arrays
    its descriptor points at."""

    def __init__(self, scn, precision):
        self.scenario = scn
        self.pk = scn.pack()
        h = C.c_void_p()
        rc = _lib.lib().af_pool_stage(C.byref(self.pk.desc), int(precision), C.byref(h))
        if rc == AF_E_INVALID:
            raise RequestError(_lib.lib().af_last_error().decode(errors="replace"))
        _lib.check("af_pool_stage", rc)
        self.handle = h

    def free(self):
        if getattr(self, "handle", None):
            _lib.lib().af_staged_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Request:
    __slots__ = ("rid", "scenario", "arrival", "admitted", "done", "steps", "cells", "result",
                 "error", "last_t", "stalls", "clock", "staged")

    def __init__(self, rid, scenario, arrival):
        self.rid, self.scenario, self.arrival = rid, scenario, arrival
        self.admitted = self.done = None
        self.steps = 0
        self.cells = scenario.cells
        self.result = None
        self.error = None
        self.last_t = None
        self.stalls = 0
        self.clock = None
        self.staged = None
        self.clock = None   # (t, steps) at the last poll


def _lane_plan(active, est, chunk_steps, ratio, cell_share):
    """The requests to put in the pool's lane, and how many steps per chunk
    the lane should take.  `est` maps an active slot to its estimated
    remaining steps.  A request joins when it needs more than `ratio` times
    the median remaining steps of the active requests (and more than two
    chunks), the longest first, while the lane holds at most `cell_share` of
    the active cells.  The lane's members leave the pool's launches, so the
    lane takes at least a chunk of steps per chunk, and as many more as let
    its longest request keep pace with the median one."""
    if len(est) < 2:
        return [], 0
    med = float(np.median(list(est.values())))
    floor = max(med * ratio, 2.0 * chunk_steps)
    cand = sorted(((rem, slot) for slot, rem in est.items() if rem > floor), reverse=True)
    budget = cell_share * sum(r.cells for r in active.values())
    lane, cells = [], 0
    for rem, slot in cand:
        if cells + active[slot].cells <= budget:
            lane.append(slot)
            cells += active[slot].cells
    if not lane:
        return [], 0
    factor = max(est[s] for s in lane) / max(med, 1.0)
    return sorted(lane), 6 * max(1, int(factor * chunk_steps / 6))


def serve(requests, pool, chunk_steps=36, clock="wall", keep_results=False, graph=True,
          stall_polls=8, lane=False, lane_ratio=3.0, lane_cells=0.1, stage_workers=4,
          stage_ahead=32):
    """Runs requests (a list of Request, sorted by arrival) through the pool.

    The pool needs admit(scenario) -> slot or None (raising RequestError for a
    scenario it refuses), release(slot), advance(steps, graph),
    poll() -> (finished slots, n_active, {active slot: (t, steps)}),
    info(slot) -> {"steps", "t", "status"}, read(slot, "H") and synchronize();
    set_lane(slots) and advance_lane(steps, graph) if it has a lane.

    clock="wall": arrivals are seconds on the wall clock from the start of the
    call (a request is admitted once it has arrived); clock="steps": arrivals
    are pool step indices (a deterministic schedule for tests).  Requests are
    admitted in arrival order.  One that an empty pool cannot take (larger
    than its HBM budget, slots or tile table) or that the pool refuses (a
    non-finite input) is rejected with `error` set rather than waited for.
    A request that stops on the device (a NaN or infinite depth, a stalled
    clock), or whose clock does not move over `stall_polls` polls in a row,
    is released with `error` set; the others keep running.

    Every pool launch steps every active request once, so a request that
    needs many more steps than the others (a small time step) would set the
    end at the pace of the whole pool's launches.  With `lane` (and a pool
    that has one, and chunks of a multiple of 6 steps), such requests move to
    the pool's lane (_lane_plan), whose small launches run on a stream of
    their own beside each chunk; the lane's steps per chunk are capped so
    that its work ends within about 80% of the chunk's time.

    With a pool that can stage (Pool.stage), the host-only part of each
    admission (validation, input conversion) runs on `stage_workers` threads
    for up to `stage_ahead` waiting requests while the pool steps; a request
    is admitted once its staging is done (on the step clock, or with nothing
    else to run, the loop waits for it).

    Returns the completed requests, with admission and completion times
    filled in, and a summary (which counts the rejected and failed ones)."""
    use_lane = lane and hasattr(pool, "set_lane") and chunk_steps % 6 == 0
    lane_slots, lane_steps, lane_cap, lane_age = [], 0, 4 * chunk_steps, 0
    stager = (concurrent.futures.ThreadPoolExecutor(stage_workers)
              if stage_workers > 0 and hasattr(pool, "stage") else None)
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
    waiting = collections.deque()
    active = {}
    done = []
    rejected = []
    failed = []
    t0 = time.perf_counter()
    step = 0
    cell_updates = 0

    def now():
        return time.perf_counter() - t0 if clock == "wall" else float(step)

    def fail(slot, r, why):
        r.error = why
        r.done = now()
        pool.release(slot)
        failed.append(r)

    while pending or waiting or active:
        n = now()
        while pending and pending[0].arrival <= n:
            waiting.append(pending.popleft())
        while waiting:
            r = waiting[0]
            try:
                if stager is not None:
                    # keep the next stage_ahead waiting requests staging
                    for k, w in enumerate(waiting):
                        if k >= stage_ahead:
                            break
                        if w.staged is None:
                            w.staged = stager.submit(pool.stage, w.scenario)
                    if not r.staged.done() and active and clock == "wall":
                        break               # admitted after the next chunk
                    item = r.staged.result()
                else:
                    item = r.scenario
                slot = pool.admit(item)
            except RequestError as e:
                waiting.popleft()           # refused for what it is: it never will fit
                r.staged = None
                r.error = f"refused: {e}"
                rejected.append(r)
                continue
            if slot is None:
                if active:
                    break                   # room frees up as active requests finish
                waiting.popleft()           # an empty pool cannot take it: it never will
                r.error = "larger than the pool (HBM budget, scenario slots or tile table)"
                rejected.append(r)
                continue
            waiting.popleft()
            r.staged = None
            r.admitted = now()
            r.last_t, r.stalls, r.clock = None, 0, None
            active[slot] = r
        if not active:
            if clock == "wall":
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
            else:
                step = int(max(step + 1, pending[0].arrival)) if pending else step + 1
            continue
        t_chunk = time.perf_counter()
        if lane_steps:
            pool.advance_lane(lane_steps, graph=graph)   # runs beside the chunk
        pool.advance(chunk_steps, graph=graph)
        step += chunk_steps
        finished, _, clocks = pool.poll()
        if lane_steps:
            # keep the lane's work within ~80% of the chunk's wall time
            wall_ms = (time.perf_counter() - t_chunk) * 1e3
            lane_ms = pool.lane_ms()
            if lane_ms > 0.9 * wall_ms:
                lane_cap = max(chunk_steps, int(lane_cap * 0.8 * wall_ms / lane_ms))
            elif lane_ms < 0.7 * wall_ms and lane_steps >= lane_cap:
                lane_cap = int(lane_cap * 1.25)
        for slot in finished:
            r = active.pop(slot)
            inf = pool.info(slot)
            r.steps = inf["steps"]
            status = inf.get("status", 0)
            if status:
                fail(slot, r, f"stopped after step {r.steps} at t = {inf.get('t', 0.0):.6g} s: "
                              f"{status_text(status)}")
                continue
            cell_updates += r.steps * r.cells
            if keep_results:
                r.result = pool.read(slot, "H")
            r.done = now()
            pool.release(slot)
            done.append(r)
        # a clock that no longer moves would hold its slot for ever
        est = {}
        for slot, r in list(active.items()):
            if slot not in clocks:
                continue
            t, n_steps = clocks[slot]
            r.stalls = r.stalls + 1 if (r.last_t is not None and not t > r.last_t) else 0
            r.last_t = t
            if r.stalls >= stall_polls:
                active.pop(slot)
                r.steps = pool.info(slot)["steps"]
                fail(slot, r, f"clock stalled at t = {t:.6g} s over {r.stalls} polls")
                continue
            if use_lane and r.clock is not None and n_steps > r.clock[1] and t > r.clock[0]:
                dt = (t - r.clock[0]) / (n_steps - r.clock[1])
                est[slot] = max(0.0, (r.scenario.sim_time - t) / dt)
            r.clock = (t, n_steps)
        if use_lane:
            plan, want = _lane_plan(active, est, chunk_steps, lane_ratio, lane_cells)
            lane_age += 1
            # a change rebuilds the tile tables: make it when a member has
            # left, or at most every fourth poll
            gone = any(s not in active for s in lane_slots)
            if plan != lane_slots and (gone or lane_age >= 4 or not lane_slots):
                pool.set_lane(plan)
                lane_slots, lane_age = plan, 0
            lane_slots = [s for s in lane_slots if s in active]
            lane_steps = (6 * max(chunk_steps // 6, min(want, lane_cap) // 6)
                          if lane_slots else 0)
    pool.synchronize()
    wall = time.perf_counter() - t0
    if stager is not None:
        stager.shutdown(wait=True)
    lat = np.array([r.done - r.arrival for r in done]) if done else np.zeros(1)
    summary = {
        "requests": len(done),
        "rejected": len(rejected),
        "failed": len(failed),
        "wall_s": wall,
        "cell_updates": cell_updates,
        "gcups": cell_updates / wall / 1e9 if clock == "wall" else None,
        "p50_latency": float(np.percentile(lat, 50)),
        "p99_latency": float(np.percentile(lat, 99)),
        "clock": clock,
    }
    return done, summary
