"""Autotuner for the fused step's tile shape and kernel variant.

For a workload (one scenario or a ragged batch), every candidate tuning
(tile rows x kernel variant x waves per block x tile order) is built, warmed up and timed
over eager launches of the step kernel with HIP events on the launch stream;
the fastest is returned with the whole table.  Results can be kept in a JSON
cache keyed by the workload's shape, so a serving process tunes each shape
class once.

Every candidate computes bitwise the same result (the tile shape and the
variant change only the work decomposition), so tuning never changes
answers; tests/test_gpu_fused.py::test_tiling_invariance checks this.
"""
import itertools
import json
import os

from .domain import Batch, PRECISIONS, _torch, make_tuning

DEFAULT_ROWS = (8, 16, 32, 64, 128)
DEFAULT_VARIANTS = (1, 2, 3)
DEFAULT_WPB = (4,)
DEFAULT_ORDERS = (0, 1)


def signature(scenarios, precision):
    """Cache key: precision, scenario count and the grid shapes."""
    shapes = sorted((s.nx, s.ny) for s in scenarios)
    cells = sum(nx * ny for nx, ny in shapes)
    return f"{precision}:{len(shapes)}:{cells}:{shapes[0]}:{shapes[-1]}"


def _time_candidate(torch, scenarios, precision, tuning, warmup, launches):
    b = Batch(scenarios, precision=precision, tuning=tuning)
    try:
        for _ in range(warmup):
            b.launch_step()
        st = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(launches):
            b.launch_step()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / launches
    finally:
        b.close()
        del b


def autotune(scenarios, precision="fp32", rows=DEFAULT_ROWS, variants=DEFAULT_VARIANTS,
             waves_per_block=DEFAULT_WPB, orders=DEFAULT_ORDERS, warmup=3, launches=20,
             cache_path=None, log=None):
    """Returns (best af_tuning, table of (ms per step, rows, variant, wpb,
    tile order)).

    `scenarios` is a list (a single domain is a list of one).  With
    cache_path, a stored result for the same signature is returned without
    timing, and a new result is stored."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision {precision!r}")
    key = signature(scenarios, precision)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            c = cache[key]
            return (make_tuning(tile_rows=c["rows"], waves_per_block=c["wpb"],
                                variant=c["variant"], tile_order=c.get("order", 0)),
                    c["table"])
    torch = _torch()
    table = []
    for r, v, w, o in itertools.product(rows, variants, waves_per_block, orders):
        tun = make_tuning(tile_rows=r, waves_per_block=w, variant=v, tile_order=o)
        ms = _time_candidate(torch, scenarios, precision, tun, warmup, launches)
        table.append((ms, r, v, w, o))
        if log:
            log(f"rows={r} variant={v} wpb={w} order={o}: {ms:.4f} ms/step")
    table.sort()
    ms, r, v, w, o = table[0]
    if cache_path:
        cache[key] = {"rows": r, "variant": v, "wpb": w, "order": o, "ms": ms, "table": table}
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return make_tuning(tile_rows=r, waves_per_block=w, variant=v, tile_order=o), table
