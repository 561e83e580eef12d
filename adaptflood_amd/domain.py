"""Fused-path runner: one or many scenarios advanced by the HIP step kernel.

Device memory for a batch is one torch uint8 tensor (the arena); the native
library lays the state out inside it and never allocates.  All compute runs
in the HIP kernels of libadaptflood_amd.so.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import AF_FP32, AF_FP64, RunInfo, Tuning, call

PRECISIONS = {"fp32": AF_FP32, "fp64": AF_FP64, AF_FP32: AF_FP32, AF_FP64: AF_FP64}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the fused path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream(torch, stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def make_tuning(tile_rows=0, waves_per_block=0, graph_steps=0, dt_log_cap=0, report_cap=0,
                variant=0, tile_order=0):
    t = Tuning()
    t.tile_rows, t.waves_per_block, t.graph_steps = tile_rows, waves_per_block, graph_steps
    t.dt_log_cap, t.report_cap, t.variant = dt_log_cap, report_cap, variant
    t.tile_order = tile_order
    return t


class Batch:
    """A batch of scenarios (a single domain is a batch of one)."""

    def __init__(self, scenarios, precision="fp32", tuning=None, device=None, stream=None):
        torch = _torch()
        self._torch = torch
        self.scenarios = list(scenarios)
        self.precision = PRECISIONS[precision]
        self.packed = [s.pack() for s in self.scenarios]
        n = len(self.packed)
        self._descs = (_lib.ScenarioDesc * n)(*[p.desc for p in self.packed])
        self.tuning = tuning if tuning is not None else make_tuning()
        nbytes = C.c_size_t(0)
        call("af_batch_arena_bytes", self._descs, n, self.precision, C.byref(self.tuning),
             C.byref(nbytes))
        self.arena_bytes = int(nbytes.value)
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.arena = torch.empty(self.arena_bytes + 256, dtype=torch.uint8, device=dev)
        base = self.arena.data_ptr()
        aligned = (base + 255) // 256 * 256
        self._stream = stream
        h = C.c_void_p()
        call("af_batch_create", self._descs, n, self.precision, C.byref(self.tuning),
             C.c_void_p(aligned), C.c_size_t(self.arena_bytes), self._s(), C.byref(h))
        self.handle = h
        ns, nt = C.c_int(), C.c_int()
        call("af_batch_size", self.handle, C.byref(ns), C.byref(nt))
        self.n_tiles = nt.value

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

    # stepping ---------------------------------------------------------------
    def advance(self, steps, graph=False):
        call("af_domain_advance", self.handle, int(steps), 1 if graph else 0, self._s())

    def launch_step(self):
        call("af_batch_launch_step", self.handle, self._s())

    def run(self, check_every=256, graph=True):
        call("af_domain_run", self.handle, int(check_every), 1 if graph else 0, self._s())

    def finished(self):
        d = C.c_int()
        call("af_batch_finished", self.handle, C.byref(d), self._s())
        return bool(d.value)

    # results ----------------------------------------------------------------
    def info(self, i=0):
        r = RunInfo()
        call("af_batch_info", self.handle, i, C.byref(r), self._s())
        return {"t": r.t, "steps": r.steps_done, "last_dt": r.last_dt,
                "n_reports": r.n_reports, "finished": bool(r.finished), "hmax": r.hmax,
                "status": r.status}

    def variant(self):
        """The kernel variant the next launches use (1..3, af_tuning.variant's
        numbering; at FP64 with variant 0 it follows the wet area)."""
        v = C.c_int32()
        call("af_domain_variant", self.handle, C.byref(v))
        return int(v.value)

    def set_sim_time(self, sim_time, i=0):
        """Moves the stop time of scenario i (runs stop at the first step with
        t >= sim_time, as the reference's loop does)."""
        call("af_batch_set_sim_time", self.handle, i, float(sim_time), self._s())

    def read(self, what="H", i=0):
        """State in the reference layout as float64: H (ny, nx); Qx, Qy,
        Qxold, Qyold ((ny+1)*(nx+1) flat, pitch nx+1)."""
        which = {"H": 0, "Qx": 1, "Qy": 2, "Qxold": 3, "Qyold": 4}[what]
        s = self.scenarios[i]
        if which == 0:
            out = np.empty((s.ny, s.nx), dtype=np.float64)
        else:
            out = np.empty((s.ny + 1) * (s.nx + 1), dtype=np.float64)
        call("af_batch_read", self.handle, i, which, _lib.dptr(out), self._s())
        return out

    def read_cells(self, idx, i=0):
        """H at flat cell indices j*nx + i of scenario i (the stage gauges of
        main.cpp:306-322), as float64; NaN for an index outside the grid."""
        idx = np.ascontiguousarray(idx, dtype=np.int64).ravel()
        out = np.empty(max(idx.size, 1), dtype=np.float64)
        call("af_batch_read_cells", self.handle, i,
             idx.ctypes.data_as(C.POINTER(C.c_int64)), int(idx.size), _lib.dptr(out), self._s())
        return out[: idx.size]

    def snapshot(self, i=0, velocity=True):
        """Device output snapshot of scenario i: H (ny, nx) and, with
        velocity, the ComputeVelocity fields Vx, Vy, as float64 numpy arrays.
        The snapshot is taken on the device behind the enqueued steps."""
        torch = _torch()
        s = self.scenarios[i]
        bufs = [torch.empty((s.ny, s.nx), dtype=torch.float64, device=self.arena.device)
                for _ in range(3 if velocity else 1)]
        ptrs = [b.data_ptr() for b in bufs] + [None] * (3 - len(bufs))
        call("af_batch_snapshot", self.handle, i, *ptrs, self._s())
        (self._stream or torch.cuda.current_stream()).synchronize()
        return [b.cpu().numpy() for b in bufs]

    def dt_log(self, i=0):
        n = C.c_int64()
        call("af_batch_read_dt_log", self.handle, i, None, 0, C.byref(n), self._s())
        out = np.empty(max(int(n.value), 1))
        call("af_batch_read_dt_log", self.handle, i, _lib.dptr(out), int(n.value), C.byref(n),
             self._s())
        return out[: int(n.value)]

    def reports(self, i=0):
        n = C.c_int32()
        call("af_batch_read_reports", self.handle, i, None, 0, C.byref(n), self._s())
        cap = int(n.value)
        out = np.empty((max(cap, 1), 4))
        call("af_batch_read_reports", self.handle, i, _lib.dptr(out), cap, C.byref(n), self._s())
        return out[: min(cap, int(n.value))]

    def volume(self, i=0):
        v, a = C.c_double(), C.c_double()
        call("af_batch_volume", self.handle, i, C.byref(v), C.byref(a), self._s())
        return v.value, a.value


class Domain(Batch):
    """One scenario on the fused path."""

    def __init__(self, scenario, precision="fp32", tuning=None, device=None, stream=None):
        super().__init__([scenario], precision, tuning, device, stream)
        self.scenario = scenario
