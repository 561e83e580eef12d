"""Benchmark: GCUPS and HBM-roofline fraction of the fused local-inertial step.

Default workload (one MI355X): the 16384 x 16384 roofline case at 2 m cells,
fractal DEM with 10% NODATA blobs, 5000 QFIX sources, all edges FREE, FP32
state, starting dry; W untimed warm-up steps, then K timed steps.  Prints one
JSON line.  The same K steps are also measured at FP64 (key "fp64"); the
served ensemble (512 requests, Poisson arrivals at 20/s, continuous batching)
reports its p50/p99 completion latency (key "served"), and the same requests
arriving at 1000/s the saturated pool's rate (key "served_saturated"); and the FP64 CPU
oracle is timed on a bounded sample of the roofline workload ("cpu_baseline").

Other workloads (--workload): "ensemble" (512 ragged scenarios in one launch
per step against sequential single-scenario runs), "served" (the served
ensemble with its sequential baselines), "accuracy" (FP32/FP64 errors against
FP64 references), "divergence" (lockstep drift at equal step counts), "shard"
(one GPU's row slab), "driver" (the rebuilt reference program end to end),
"lone" (step time of small and mid-size domains run alone), "twin" (the
FP64 path against itself with inputs nudged by one ulp, beside FP32),
"ceiling" (the bandwidth of the step's memory pattern without the step),
"configs" (the plane 256^2 and valley 4096^2 workloads at FP32 and FP64
beside the CPU oracle) and "budget" (on the CPU: the volume budget of the
FP64 twin pair, sources, edge flow and depth-clamp gains, with the oracle).

Algorithmic bytes per cell-update: read z, H, Qx_old, Qy_old and write H, Qx,
Qy = 7 values = 28 B at FP32 (56 B at FP64).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS + % HBM roofline (FP32 vs FP64); p50/p99 scenario latency"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_step_roofline.md")


def _traffic_from_profile(precision):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), as committed in
    profiles/.  Returns None when no measurement is recorded."""
    try:
        with open(PROFILE_TRAFFIC) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch_{precision}:"):
                    return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def host_cpu():
    """The host's CPU model and physical core count (lscpu), and nproc."""
    import subprocess
    model, cores, sockets = "unknown", None, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Core(s) per socket":
                cores = int(v)
            elif k.strip() == "Socket(s)":
                sockets = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    phys = cores * sockets if cores and sockets else None
    return {"model": model, "physical_cores": phys, "nproc": os.cpu_count()}


def measure(scn, precision, steps, warmup, tile_rows, wpb, variant=0):
    import torch
    from adaptflood_amd.domain import Domain, make_tuning
    d = Domain(scn, precision=precision,
               tuning=make_tuning(tile_rows=tile_rows, waves_per_block=wpb, graph_steps=36,
                                  variant=variant))
    torch.cuda.synchronize()
    d.advance(warmup, graph=False)
    torch.cuda.synchronize()
    # timed region: exactly `steps` steps, replayed from captured graphs
    d.advance(0, graph=True)  # capture outside the timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-launch kernel time with HIP events on the launch stream
    n_ev = min(steps, 50)
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(n_ev):
        d.launch_step()
    e1.record(st)
    torch.cuda.synchronize()
    kern_s = e0.elapsed_time(e1) / 1e3 / n_ev
    info = d.info()
    d.close()
    return wall, kern_s, info


def _ensemble_member(k, seed_base, sim_time):
    from adaptflood_amd import generators
    return generators.ensemble_member(k, seed_base=seed_base, sim_time=sim_time)


def generate_members(ks, seed_base, sim_times, workers=8):
    """Ensemble members in worker processes (forked before any GPU use)."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_ensemble_member, ks, [seed_base] * len(ks), sim_times, chunksize=4))


def _event_ms(torch, launch, n):
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(n):
        launch()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def sequential_fused(scns, precision, steps, warmup=3):
    """Each scenario alone through the fused single-domain path (graph-replayed),
    back to back: aggregate cell-updates/s over the timed steps."""
    import torch
    from adaptflood_amd.domain import Domain, make_tuning
    cu, sec = 0, 0.0
    for scn in scns:
        d = Domain(scn, precision=precision, tuning=make_tuning(graph_steps=36))
        d.advance(warmup, graph=False)
        d.advance(0, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.advance(steps, graph=True)
        torch.cuda.synchronize()
        sec += time.perf_counter() - t0
        cu += scn.cells * steps
        d.close()
    return cu / sec / 1e9, sec


def sequential_fused_own(scns, own_steps, precision, steps, warmup=3):
    """The same single-scenario timing as sequential_fused, weighted two ways:
    uniform (every scenario `steps` steps, as sequential_fused) and by each
    scenario's own step count (`own_steps`, e.g. what it took in a served
    run), i.e. the time to run the scenarios to completion one after another
    at their measured lone step times."""
    import torch
    from adaptflood_amd.domain import Domain, make_tuning
    cu, sec, cu_own, sec_own = 0, 0.0, 0, 0.0
    for scn, own in zip(scns, own_steps):
        d = Domain(scn, precision=precision, tuning=make_tuning(graph_steps=36))
        d.advance(warmup, graph=False)
        d.advance(0, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.advance(steps, graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        d.close()
        sec += dt
        cu += scn.cells * steps
        sec_own += dt / steps * own
        cu_own += scn.cells * own
    return cu / sec / 1e9, cu_own / sec_own / 1e9, sec_own


def sequential_plain(scns, steps):
    """Each scenario alone through the reference's call sequence (the compat
    C ABI, FP64, one synchronous call per kernel, dt read back every step)."""
    from adaptflood_amd import kernels
    cu, sec = 0, 0.0
    for scn in scns:
        r = kernels.run_sequence(scn, max_steps=steps)
        sec += r["loop_seconds"]
        cu += scn.cells * r["steps"]
    return cu / sec / 1e9, sec


def bench_ensemble(args):
    """512 ragged scenarios (sides log-uniform 128..2048, aspect 1..4) advanced
    together, one launch per step, against the same scenarios run back to back
    through the single-scenario paths."""
    t = time.perf_counter()
    n = args.members
    scns = generate_members(list(range(n)), 4000, [3600.0] * n)
    cells = sum(s.cells for s in scns)
    print(f"[bench] generated {n} scenarios, {cells} cells, in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    import torch
    from adaptflood_amd.domain import Batch, make_tuning
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    prec = args.precision
    tun = make_tuning(graph_steps=36)
    tuned = None
    if args.autotune:
        from adaptflood_amd.tune import autotune
        t = time.perf_counter()
        best, table = autotune(scns, prec, rows=(16, 32, 64, 128), launches=10,
                               log=lambda m: print(f"[tune] {m}", file=sys.stderr, flush=True))
        tun = make_tuning(tile_rows=best.tile_rows, waves_per_block=best.waves_per_block,
                          variant=best.variant, graph_steps=36, tile_order=best.tile_order)
        tuned = {"tile_rows": best.tile_rows, "variant": best.variant,
                 "waves_per_block": best.waves_per_block, "tile_order": best.tile_order,
                 "tuning_s": round(time.perf_counter() - t, 1)}
        print(f"[bench] autotuned: {tuned}", file=sys.stderr, flush=True)
    b = Batch(scns, precision=prec, tuning=tun)
    b.advance(args.warmup, graph=False)
    b.advance(0, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(args.steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = _event_ms(torch, b.launch_step, 20)
    total = args.warmup + args.steps + 20
    active = sum(1 for k in range(n) if b.info(k)["steps"] == total)
    gcups = cells * args.steps / wall / 1e9
    B = 28.0 if prec == "fp32" else 56.0
    achieved = B * cells / (kern_ms * 1e-3) / 1e9
    print(f"[bench] batched {prec}: {wall * 1e3 / args.steps:.3f} ms/step, {gcups:.1f} GCUPS, "
          f"kernel {kern_ms:.3f} ms = {achieved:.0f} GB/s, {active}/{n} active",
          file=sys.stderr, flush=True)
    b.close()
    fused_sample = scns[::args.seq_stride]
    g_fused, s_fused = sequential_fused(fused_sample, prec, args.seq_steps)
    print(f"[bench] sequential fused {prec} ({len(fused_sample)} scenarios): {g_fused:.2f} GCUPS",
          file=sys.stderr, flush=True)
    plain_sample = scns[::args.plain_stride]
    g_plain, s_plain = sequential_plain(plain_sample, args.plain_steps)
    print(f"[bench] sequential plain fp64 ({len(plain_sample)} scenarios): {g_plain:.2f} GCUPS",
          file=sys.stderr, flush=True)
    out = {
        "metric": "aggregate cell-updates/s, batched ensemble vs sequential single-scenario runs",
        "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"ensemble of {n} ragged scenarios (sides log-uniform 128..2048, "
                               f"aspect U(1,4), dx 5 m, fractal DEM, 10-500 QVAR, rain, "
                               f"own adaptive dt), one launch per step, graph-replayed",
                   "scenarios": n, "cells": cells, "active_in_timed_steps": active,
                   "precision": prec, "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(kern_ms, 4)},
        "sequential_fused": {"value": round(g_fused, 3), "unit": "GCUPS",
                             "sample": f"every {args.seq_stride}th scenario "
                                       f"({len(fused_sample)}), {args.seq_steps} steps each, "
                                       f"{prec}, graph-replayed", "seconds": round(s_fused, 3)},
        "sequential_plain": {"value": round(g_plain, 3), "unit": "GCUPS",
                             "sample": f"every {args.plain_stride}th scenario "
                                       f"({len(plain_sample)}), {args.plain_steps} steps each, "
                                       f"fp64 reference call sequence", "seconds": round(s_plain, 3)},
        "speedup_vs_sequential_fused": round(gcups / g_fused, 2),
        "speedup_vs_sequential_plain": round(gcups / g_plain, 2),
    }
    print(json.dumps(out), flush=True)


def run_served(scns, arrivals, prec, budget_gb, chunk, lane=False, variant=0):
    """The continuous-batching loop over requests arriving at `arrivals` (s)."""
    import torch
    from adaptflood_amd import generators
    from adaptflood_amd.domain import make_tuning
    from adaptflood_amd.serve import Pool, Request, serve
    n = len(scns)
    free, _ = torch.cuda.mem_get_info()
    budget = min(int(budget_gb * (1 << 30)), int(free * 0.8))
    pool = Pool(budget_bytes=budget, precision=prec, max_scenarios=n, max_tiles=1 << 20,
                tuning=make_tuning(variant=variant))
    # a few requests through a warm pool first, so kernel loading is not timed
    warm = [Request(-1 - k, generators.ensemble_member(k, seed_base=5000, sim_time=400.0), 0.0)
            for k in range(2)]
    serve(warm, pool, chunk_steps=36, clock="wall")
    reqs = [Request(k, s, float(a)) for k, (s, a) in enumerate(zip(scns, arrivals))]
    done, summary = serve(reqs, pool, chunk_steps=chunk, clock="wall", lane=lane)
    print(f"[bench] served {summary['requests']} in {summary['wall_s']:.1f}s: "
          f"{summary['gcups']:.1f} GCUPS, p50 {summary['p50_latency']:.2f}s, "
          f"p99 {summary['p99_latency']:.2f}s", file=sys.stderr, flush=True)
    pool.close()
    del pool
    torch.cuda.empty_cache()
    return done, summary, budget


def bench_served(args):
    """Poisson arrivals (20/s) of ensemble scenarios with sim_time U(600, 3600) s,
    admitted into a pool by HBM footprint, advanced together, released when done."""
    from adaptflood_amd import generators
    n = args.members
    arr, sims = generators.arrivals(n, rate=args.rate, seed=5)
    t = time.perf_counter()
    scns = generate_members(list(range(n)), 5000, [float(x) for x in sims])
    cells = sum(s.cells for s in scns)
    print(f"[bench] generated {n} requests, {cells} cells, in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    prec = args.precision
    done, summary, budget = run_served(scns, arr, prec, args.budget_gb, args.chunk,
                                       lane=args.lane, variant=args.variant)
    own = {r.rid: r.steps for r in done}
    pick = [k for k in range(0, n, args.seq_stride) if k in own]
    g_fused, g_own, s_own = sequential_fused_own([scns[k] for k in pick], [own[k] for k in pick],
                                                 prec, args.seq_steps)
    print(f"[bench] sequential fused {prec} ({len(pick)} scenarios): {g_fused:.2f} GCUPS "
          f"({args.seq_steps} steps each), {g_own:.2f} GCUPS (own step counts)",
          file=sys.stderr, flush=True)
    plain_sample = scns[::args.plain_stride]
    g_plain, s_plain = sequential_plain(plain_sample, args.plain_steps)
    print(f"[bench] sequential plain fp64 ({len(plain_sample)} scenarios): {g_plain:.2f} GCUPS",
          file=sys.stderr, flush=True)
    steps = np.array([r.steps for r in done])
    out = {
        "metric": "served ensemble: p50/p99 scenario latency and aggregate cell-updates/s",
        "value": round(summary["gcups"], 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": int(steps.sum()), "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"{n} requests, Poisson arrivals at {args.rate}/s, ensemble "
                               f"generator (seed 5), sim_time U(600,3600) s, HBM budget "
                               f"{budget / (1 << 30):.0f} GiB, chunks of {args.chunk} steps"
                               f"{', with the lane' if args.lane else ''}",
                   "requests": n, "cells": cells, "precision": prec},
        "p50_latency_s": round(summary["p50_latency"], 3),
        "p99_latency_s": round(summary["p99_latency"], 3),
        "wall_s": round(summary["wall_s"], 3),
        "arrival_span_s": round(float(arr[-1]), 3),
        "cell_updates": int(summary["cell_updates"]),
        "sequential_fused": {"value": round(g_fused, 3), "unit": "GCUPS",
                             "sample": f"every {args.seq_stride}th request ({len(pick)}), "
                                       f"{args.seq_steps} steps each"},
        "sequential_fused_own_steps": {
            "value": round(g_own, 3), "unit": "GCUPS",
            "sample": "the same requests at their measured lone step times, each for the steps "
                      "it took when served", "seconds": round(s_own, 3)},
        "sequential_plain": {"value": round(g_plain, 3), "unit": "GCUPS",
                             "sample": f"every {args.plain_stride}th request "
                                       f"({len(plain_sample)}), {args.plain_steps} steps each, "
                                       f"fp64 reference call sequence"},
        "speedup_vs_sequential_fused": round(summary["gcups"] / g_fused, 2),
        "speedup_vs_sequential_fused_own_steps": round(summary["gcups"] / g_own, 2),
        "speedup_vs_sequential_plain": round(summary["gcups"] / g_plain, 2),
    }
    print(json.dumps(out), flush=True)


def _errors(H, ref, dA, wet=0.01):
    """max and RMS |dH|, inundated-area difference (cells deeper than `wet`,
    the reference's Area criterion, main.cpp:296) and relative mass error."""
    d = np.abs(H - ref)
    a, b = H > wet, ref > wet
    vol, vref = float(np.sum(H)) * dA, float(np.sum(ref)) * dA
    return {"max_abs_dH_m": float(d.max()), "rms_dH_m": float(np.sqrt(np.mean(d * d))),
            "area_m2": float(a.sum()) * dA, "area_ref_m2": float(b.sum()) * dA,
            "area_diff_m2": float(a.sum() - b.sum()) * dA,
            "area_mismatch_cells": int(np.count_nonzero(a != b)),
            "rel_mass_err": abs(vol - vref) / vref if vref else 0.0}


def _full_run(scn, precision, label):
    """Runs to sim_time in graph chunks, printing progress."""
    import torch
    from adaptflood_amd.domain import Domain, make_tuning
    d = Domain(scn, precision=precision, tuning=make_tuning(graph_steps=216))
    torch.cuda.synchronize()
    t0 = last = time.perf_counter()
    while True:
        d.advance(2160, graph=True)
        info = d.info()
        if info["finished"]:
            break
        if time.perf_counter() - last > 20:
            last = time.perf_counter()
            print(f"[acc] {label} {precision}: t={info['t']:.0f}s steps={info['steps']} "
                  f"({last - t0:.0f}s)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    H = d.read("H")
    d.close()
    return H, info, wall


def bench_accuracy(args):
    """FP32 and FP64 against FP64 references.  The 4096^2 valley case to
    --acc-time against the FP64 CPU oracle, and to its full 7200 s at FP32
    against the FP64 GPU path (itself pinned to the oracle on the smaller
    runs); the 16384^2 roofline case to --roof-time (its full 3600 s by
    default) at FP32 against the FP64 GPU path."""
    import torch
    cases = {c.strip() for c in args.acc_cases.split(",") if c.strip()}
    lines = []
    if "valley" in cases:
        lines += _accuracy_valley(args)
    if "roofline" in cases and args.size > 0:
        lines += _accuracy_roofline(args)
    torch.cuda.synchronize()
    print(json.dumps({"metric": "FP32/FP64 error against FP64 reference paths",
                      "results": lines}), flush=True)


def _accuracy_valley(args):
    from adaptflood_amd import generators
    from oracle import oracle
    lines = []
    t = time.perf_counter()
    scn = generators.valley_case(n=args.valley_size)
    full_time = scn.sim_time
    scn.sim_time = min(args.acc_time, full_time)
    print(f"[acc] valley {scn.nx}x{scn.ny} generated in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    dA = scn.dx * scn.dx
    t = time.perf_counter()
    run = oracle.OracleRun(scn, dt_log_cap=1, mass_log_cap=1)
    while run.t < scn.sim_time:
        if run.advance(500) == 0:
            break
        print(f"[acc] oracle t={run.t:.0f}s steps={run.steps} ({time.perf_counter() - t:.0f}s)",
              file=sys.stderr, flush=True)
    ref = run.state()
    t_orc = time.perf_counter() - t
    for prec in ("fp64", "fp32"):
        H, info, wall = _full_run(scn, prec, "valley")
        e = _errors(H, ref["H"], dA)
        e.update({"case": f"valley {scn.nx}x{scn.ny}, {scn.sim_time:.0f} s", "precision": prec,
                  "against": "FP64 CPU oracle", "steps": info["steps"],
                  "steps_ref": ref["steps"], "gpu_wall_s": round(wall, 3),
                  "oracle_wall_s": round(t_orc, 1), "oracle_threads": oracle.num_threads()})
        print(f"[acc] {json.dumps(e)}", file=sys.stderr, flush=True)
        lines.append(e)
    del ref, run
    scn.sim_time = full_time
    H64, i64, w64 = _full_run(scn, "fp64", "valley")
    H32, i32, w32 = _full_run(scn, "fp32", "valley")
    e = _errors(H32, H64, dA)
    e.update({"case": f"valley {scn.nx}x{scn.ny}, {full_time:.0f} s", "precision": "fp32",
              "against": "FP64 GPU path", "steps": i32["steps"], "steps_ref": i64["steps"],
              "gpu_wall_s": round(w32, 3), "ref_wall_s": round(w64, 3)})
    print(f"[acc] {json.dumps(e)}", file=sys.stderr, flush=True)
    lines.append(e)
    return lines


def _accuracy_roofline(args):
    from adaptflood_amd import generators
    t = time.perf_counter()
    big = generators.roofline_case(n=args.size)
    big.sim_time = args.roof_time
    print(f"[acc] roofline {big.nx}x{big.ny} generated in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    H64, i64, w64 = _full_run(big, "fp64", "roofline")
    H32, i32, w32 = _full_run(big, "fp32", "roofline")
    e = _errors(H32, H64, big.dx * big.dx)
    e.update({"case": f"roofline {big.nx}x{big.ny}, {big.sim_time:.0f} s",
              "precision": "fp32", "against": "FP64 GPU path", "steps": i32["steps"],
              "steps_ref": i64["steps"], "gpu_wall_s": round(w32, 3),
              "ref_wall_s": round(w64, 3)})
    print(f"[acc] {json.dumps(e)}", file=sys.stderr, flush=True)
    return [e]


def bench_divergence(args):
    """How far two runs of the same case drift apart at equal step counts:
    the FP64 fused path against the CPU oracle, and FP32 against FP64, on the
    4096^2 valley case (an inflow the scheme amplifies) and an 8192^2 roofline
    case.  Prints max |dH| every `every` steps."""
    import torch
    from adaptflood_amd import generators
    from adaptflood_amd.domain import Domain, make_tuning
    from oracle import oracle
    out = {"metric": "max |dH| between runs at equal step counts", "series": []}
    cases_on = {c.strip() for c in args.div_cases.split(",") if c.strip()}

    def record(name, rows):
        out["series"].append({"case": name, "points": rows})
        print(f"[div] {name}: " + ", ".join(f"{k}:{v:.2e}" for k, v, *_ in rows),
              file=sys.stderr, flush=True)

    cases = []
    if "valley" in cases_on:
        valley = generators.valley_case(n=args.valley_size)
        r = oracle.OracleRun(valley, dt_log_cap=1, mass_log_cap=1)
        d = Domain(valley, precision="fp64", tuning=make_tuning())
        rows, done = [], 0
        while done < 600:
            r.advance(100)
            d.advance(100)
            done += 100
            diff = float(np.max(np.abs(d.read("H").ravel() - r.H)))
            rows.append((done, diff, None, None))
        record(f"valley {valley.nx}^2: FP64 fused vs FP64 CPU oracle", rows)
        d.close()
        del r
        cases.append(("valley", valley, 400, 40))
    if "roofline" in cases_on:
        cases.append(("roofline", generators.roofline_case(n=8192), 1200, 200))
    for name, scn, steps, every in cases:
        a = Domain(scn, precision="fp64", tuning=make_tuning())
        b = Domain(scn, precision="fp32", tuning=make_tuning())
        rows, done = [], 0
        while done < steps:
            a.advance(every)
            b.advance(every)
            done += every
            Ha, Hb = a.read("H"), b.read("H")
            rms = float(np.sqrt(np.mean((Ha - Hb) ** 2)))
            va, vb = a.volume()[0], b.volume()[0]
            rows.append((done, float(np.max(np.abs(Ha - Hb))), rms,
                         abs(vb - va) / va if va else 0.0))
        record(f"{name} {scn.nx}^2: FP32 vs FP64 fused (step, max, RMS, rel. volume diff)",
               rows)
        a.close()
        b.close()
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


def _nudged(scn, how="source"):
    """The same case with inputs moved by one ulp.  how="source": one input,
    the first QFIX value, or else the peak of the first QVAR hydrograph (a
    perturbation at one place).  how="dem": every valid terrain value (a
    perturbation of every face at FP64's rounding level, as FP32 rounding
    perturbs every cell)."""
    import dataclasses
    from adaptflood_amd._lib import AF_BC_QFIX, AF_BC_QVAR
    if how == "dem":
        dem = np.array(scn.dem, dtype=np.float64)
        valid = dem != scn.nodata
        dem[valid] = np.nextafter(dem[valid], np.inf)
        return dataclasses.replace(scn, dem=dem), "every valid DEM value + 1 ulp"
    pts = [dataclasses.replace(p) for p in scn.points]
    for p in pts:
        if p.type == AF_BC_QFIX:
            p.value = float(np.nextafter(p.value, np.inf))
            return dataclasses.replace(scn, points=pts), "first QFIX value + 1 ulp"
    for p in pts:
        if p.type == AF_BC_QVAR:
            v = np.array(p.values, dtype=np.float64)
            k = int(np.argmax(v))
            v[k] = np.nextafter(v[k], np.inf)
            p.values = v
            return dataclasses.replace(scn, points=pts), "first QVAR peak + 1 ulp"
    raise ValueError("no source to nudge")


def bench_twin(args):
    """FP64 twin: the FP64 fused path on a case and on the same case with one
    source nudged by one ulp, beside the FP32 path, in lockstep at equal step
    counts (then each to its end).  If the FP64 twin drifts from the FP64 run
    as far as FP32 does, the FP32 gap is the flow amplifying any difference,
    not an FP32 accounting error; the longest window in which the twin stays
    within 1e-9 m is where a non-bitwise path can be held to the bars.
    --twin-cases: comma-separated case:size:sim_time (case valley or roofline)."""
    import torch
    from adaptflood_amd import generators
    from adaptflood_amd.domain import Domain, make_tuning
    out = {"metric": "FP64 twin (one input +1 ulp) and FP32 against FP64, equal step counts",
           "cases": []}
    for spec in args.twin_cases.split(","):
        name, size, sim = spec.split(":")
        size, sim = int(size), float(sim)
        t = time.perf_counter()
        base = (generators.valley_case(n=size) if name == "valley"
                else generators.roofline_case(n=size))
        base.sim_time = sim
        twin, how = _nudged(base, args.twin_nudge)
        print(f"[twin] {name} {size}^2 generated in {time.perf_counter() - t:.1f}s ({how})",
              file=sys.stderr, flush=True)
        dA = base.dx * base.dx
        tun = make_tuning(graph_steps=36)
        a = Domain(base, precision="fp64", tuning=tun)
        b = Domain(twin, precision="fp64", tuning=tun)
        c = Domain(base, precision="fp32", tuning=tun)
        rows, done, nxt = [], 0, 25
        t0 = time.perf_counter()

        def compare(label):
            Ha = a.read("H")
            Hb, Hc = b.read("H"), c.read("H")
            ia, ib, ic = a.info(), b.info(), c.info()
            eb, ec = _errors(Hb, Ha, dA), _errors(Hc, Ha, dA)
            row = {"at": label, "steps": [ia["steps"], ib["steps"], ic["steps"]],
                   "t": ia["t"],
                   "twin_max": eb["max_abs_dH_m"], "twin_rms": eb["rms_dH_m"],
                   "twin_mass": eb["rel_mass_err"], "twin_cells": eb["area_mismatch_cells"],
                   "fp32_max": ec["max_abs_dH_m"], "fp32_rms": ec["rms_dH_m"],
                   "fp32_mass": ec["rel_mass_err"], "fp32_cells": ec["area_mismatch_cells"]}
            rows.append(row)
            print(f"[twin] {name} {size}^2 {label} t={ia['t']:.1f}: twin max {row['twin_max']:.2e} "
                  f"rms {row['twin_rms']:.2e} mass {row['twin_mass']:.2e} | fp32 max "
                  f"{row['fp32_max']:.2e} rms {row['fp32_rms']:.2e} mass {row['fp32_mass']:.2e} "
                  f"({time.perf_counter() - t0:.0f}s)", file=sys.stderr, flush=True)
            return ia, ib, ic

        while True:
            for d in (a, b, c):
                d.advance(nxt - done, graph=True)
            done = nxt
            ia, ib, ic = compare(f"step {done}")
            if ia["finished"] or ib["finished"] or ic["finished"]:
                break
            nxt = done * 2 if done < args.twin_linear else done + args.twin_linear
        # each run to its own end, then the final states
        for d in (a, b, c):
            while not d.info()["finished"]:
                d.advance(2160, graph=True)
        compare("end")
        window = 0
        for r in rows:
            if r["at"].startswith("step") and r["twin_max"] <= 1e-9:
                window = r["steps"][0]
        out["cases"].append({"case": f"{name} {size}x{size}, {sim:.0f} s", "nudge": how,
                             "twin_within_1e-9_m_through_step": window, "rows": rows})
        for d in (a, b, c):
            d.close()
        del a, b, c
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def bench_budget(args):
    """Volume budget of the FP64 twin pair, on the CPU: the FP64 oracle runs
    the roofline generator at --budget-size and its twin (every valid terrain
    value + 1 ulp) for --budget-steps steps each, with the build that counts
    the depth its clamps at zero add.  Each step's volume change splits into
    the QFIX inflow (v*dx*dt per source, main.cpp:246-253), the net inflow
    through the edge faces (their Q after BCs, times dt) and the clamp gains
    (kernels.cpp:287); the difference between the two runs' volumes then
    splits into the differences of those parts."""
    from adaptflood_amd import generators
    from oracle import oracle
    kw = {"n_qfix": args.budget_qfix} if args.budget_qfix > 0 else {}
    base = generators.roofline_case(n=args.budget_size, **kw)
    twin, how = _nudged(base, "dem")
    nx, ny, dA = base.nx, base.ny, base.dx * base.dx
    qfix = sum(p.value for p in base.points)
    L = oracle.lib(clamp=True)
    marks = []
    m = 25
    while m < args.budget_steps:
        marks.append(m)
        m *= 2
    marks.append(args.budget_steps)

    def edge_inflow(r):
        qx = r.Qx.reshape(-1, nx + 1)
        qy = r.Qy.reshape(-1, nx + 1)
        return (float(qx[:ny, 0].sum()) - float(qx[:ny, nx].sum()) + float(qy[0, :nx].sum())
                - float(qy[ny, :nx].sum()))

    runs = {}
    for name, scn in (("base", base), ("twin", twin)):
        r = oracle.OracleRun(scn, dt_log_cap=1, mass_log_cap=1, clamp=True)
        L.orc_clamp_gain(1)
        v0 = float(r.H.sum()) * dA
        acc = {"src": 0.0, "edge": 0.0, "clamp": 0.0}
        rows = []
        t0 = time.perf_counter()
        while r.steps < args.budget_steps and r.t < scn.sim_time:
            t_prev = r.t
            r.advance(1)
            dt = r.t - t_prev
            acc["src"] += qfix * base.dx * dt
            acc["edge"] += edge_inflow(r) * dt
            acc["clamp"] += L.orc_clamp_gain(1) * dA
            if r.steps in marks:
                v = float(r.H.sum()) * dA
                rows.append({"steps": r.steps, "t": r.t, "V": v, **acc,
                             "residual": (v - v0) - acc["src"] - acc["edge"] - acc["clamp"]})
                print(f"[budget] {name} step {r.steps} t={r.t:.1f}: V={v:.6e} src={acc['src']:.6e} "
                      f"edge={acc['edge']:.6e} clamp={acc['clamp']:.6e} residual "
                      f"{rows[-1]['residual']:.2e} ({time.perf_counter() - t0:.0f}s)",
                      file=sys.stderr, flush=True)
        runs[name] = rows
    diff = []
    for a, b in zip(runs["base"], runs["twin"]):
        d = {"steps": a["steps"], "rel_volume_diff": (b["V"] - a["V"]) / a["V"]}
        for k in ("src", "edge", "clamp"):
            d["d_" + k] = b[k] - a[k]
        diff.append(d)
        print(f"[budget] step {a['steps']}: twin - base: V {b['V'] - a['V']:.4e} "
              f"({d['rel_volume_diff']:.2e}) = src {d['d_src']:.4e} + edge {d['d_edge']:.4e} "
              f"+ clamp {d['d_clamp']:.4e}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "volume budget of the FP64 twin pair (CPU oracle)",
                      "config": {"workload": f"roofline generator {nx}x{ny}, {len(base.points)} "
                                             f"QFIX, free edges; twin: {how}",
                                 "steps": args.budget_steps},
                      "base": runs["base"], "twin": runs["twin"], "difference": diff}),
          flush=True)


def _asc_body(path):
    """The grid lines of an ESRI-ASCII file (the six header lines dropped)."""
    with open(path, "rb") as f:
        data = f.read()
    cut = 0
    for _ in range(6):
        cut = data.index(b"\n", cut) + 1
    return data[cut:]


def bench_driver(args):
    """lisflood_amd, the reference program rebuilt on this library, end to
    end on the 4096^2 valley case written in the reference's input formats:
    parse, upload, step to --drv-time, save every --save-int s with
    velocities, write the final grids.  It runs once with the asynchronous
    save path (the reference's default) and once with the synchronous one,
    and checks that both leave the same grids."""
    import re
    import shutil
    import subprocess
    import tempfile
    from adaptflood_amd import formats, generators
    exe = os.path.join(ROOT, "adaptflood_amd", "_lib", "lisflood_amd")
    if not os.path.exists(exe):
        print("lisflood_amd is not built (run __graft_entry__.build())", file=sys.stderr)
        sys.exit(1)
    scn = generators.valley_case(n=args.valley_size)
    scn.sim_time = args.drv_time
    tmp = tempfile.mkdtemp(prefix="af_driver_")
    try:
        t = time.perf_counter()
        base = formats.write_case(os.path.join(tmp, "case"), scn)
        print(f"[driver] {scn.nx}x{scn.ny} case written in {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
        with open(base) as f:
            text = f.read()
        runs = {}
        for mode in (1, 0):
            out = os.path.join(tmp, f"out{mode}")
            par = os.path.join(tmp, f"run{mode}.par")
            with open(par, "w") as f:   # later keys override the case's own
                f.write(text + f"dirroot {out}\nsaveint {args.save_int!r}\nvoutput\n"
                        f"async_output {mode}\n")
            t0 = time.perf_counter()
            r = subprocess.run([exe, "--path", args.precision, "--quiet", par],
                               capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                print(r.stdout[-2000:], r.stderr[-2000:], file=sys.stderr)
                sys.exit(1)
            loop = float(re.search(r"Wall clock time: ([\d.]+) s", r.stdout).group(1))
            steps = int(re.search(r"Steps: (\d+)", r.stdout).group(1))
            runs[mode] = {"process_s": round(wall, 2), "loop_s": loop, "steps": steps,
                          "loop_gcups": round(scn.cells * steps / loop / 1e9, 2),
                          "files": sorted(os.listdir(out)), "dir": out}
            print(f"[driver] async_output {mode}: process {wall:.1f}s, loop {loop:.2f}s, "
                  f"{steps} steps, {len(runs[mode]['files'])} files", file=sys.stderr, flush=True)
        a, b = runs[1], runs[0]
        grids = [f for f in b["files"] if f.endswith(".asc")]
        same = all(_asc_body(os.path.join(a["dir"], f)) == _asc_body(os.path.join(b["dir"], f))
                   for f in grids)
        extra = sorted(set(a["files"]) - set(b["files"]))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    for r in runs.values():
        del r["files"], r["dir"]
    print(json.dumps({
        "metric": "lisflood_amd end to end (parse, step, saves, final grids)",
        "value": runs[1]["process_s"], "unit": "s", "higher_is_better": False,
        "config": {"workload": f"valley {scn.nx}x{scn.ny} at {scn.dx:g} m, {len(scn.points)} "
                               f"points, rain; sim_time {scn.sim_time:g} s, saves every "
                               f"{args.save_int:g} s with velocity",
                   "precision": args.precision},
        "async": runs[1], "sync": runs[0],
        "grids_equal_across_modes": same, "grids_compared": len(grids),
        "async_only_files": len(extra)}), flush=True)


def bench_lone(args):
    """Step time of lone small and mid-size domains with the default tiling
    (graph-replayed): the latency-bound regime of a single scenario, a served
    request running by itself, or the sequential baselines.  Domains: wet flat
    grids and members of the served ensemble generator (seed 5000+k)."""
    import ctypes as C
    from adaptflood_amd import generators
    from adaptflood_amd._lib import call
    from adaptflood_amd.scenario import Scenario

    def flat(nx, ny):
        dem = np.tile(np.linspace(2.0, 0.0, nx), (ny, 1))
        return Scenario(dem=dem, dx=5.0, sim_time=1e9, h0=np.full((ny, nx), 0.5),
                        name=f"flat {nx}x{ny}, wet")
    cases = [flat(128, 32), flat(2048, 512)]
    for k in (0, 7, 33):
        m = generators.ensemble_member(k, seed_base=5000, sim_time=1e9)
        m.name = f"ensemble member {k}"
        cases.append(m)
    import torch
    from adaptflood_amd.domain import Domain
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    rows = []
    for prec in (("fp32",) if args.no_fp64 else ("fp32", "fp64")):
        W = 256 if prec == "fp32" else 128
        for scn in cases:
            d = Domain(scn, precision=prec)
            tiles = C.c_int32()
            call("af_pool_stats", d.handle, None, C.byref(tiles), None, None)
            strips = (scn.nx + W - 1) // W
            R = next(r for r in (64, 32, 16, 8, 4, 2, 1)
                     if strips * ((scn.ny + r - 1) // r) == tiles.value)
            d.advance(72, graph=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d.advance(args.lone_steps, graph=True)
            torch.cuda.synchronize()
            us = (time.perf_counter() - t0) / args.lone_steps * 1e6
            d.close()
            row = {"domain": scn.name, "nx": scn.nx, "ny": scn.ny, "points": len(scn.points),
                   "precision": prec, "tiles": tiles.value, "tile_rows": R,
                   "us_per_step": round(us, 2)}
            rows.append(row)
            print(f"[bench] lone {row}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "lone-domain step time (default tiling)", "unit": "us",
                      "higher_is_better": False, "steps": args.lone_steps,
                      "config": {"workload": "wet flat grids and served-ensemble members, "
                                             "each alone, graph-replayed steps"},
                      "domains": rows}), flush=True)


def bench_configs(args):
    """The two smaller single-domain workloads, each at FP32 and FP64 beside
    the FP64 CPU oracle on the host's cores:
      plane 256^2 at 10 m (one QFIX, FREE east edge), the full 3600 s run on
        the GPU (graph-replayed) and on the CPU;
      valley 4096^2 at 5 m (2000 QVAR, rain, HFIX west edge), `--steps`
        graph-replayed steps after `--warmup` on the GPU (from the dry
        start), and on the CPU steps 11-40 after 10 untimed steps.
    GCUPS count every cell of the grid; the roofline fraction uses the
    step's algorithmic bytes (28 B FP32, 56 B FP64 per cell) over the
    event-timed launch."""
    import torch
    from adaptflood_amd import generators
    from adaptflood_amd.domain import Domain, make_tuning
    from oracle import oracle
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    host = host_cpu()
    out = {"metric": "GCUPS and HBM roofline fraction, single-domain workloads",
           "cpu": {"threads": oracle.num_threads(), "host_cpu": host["model"],
                   "host_physical_cores": host["physical_cores"]}, "rows": []}
    plane = generators.plane_case()
    for prec in ("fp32", "fp64"):
        d = Domain(plane, precision=prec, tuning=make_tuning(graph_steps=216))
        d.advance(0, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.run(check_every=2160, graph=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        steps = d.info()["steps"]
        d.close()
        row = {"workload": "plane 256x256, full 3600 s", "precision": prec, "steps": steps,
               "seconds": round(wall, 4), "gcups": round(plane.cells * steps / wall / 1e9, 3)}
        out["rows"].append(row)
        print(f"[bench] configs {row}", file=sys.stderr, flush=True)
    r = oracle.OracleRun(plane, dt_log_cap=1 << 12, mass_log_cap=64)
    r.advance()
    row = {"workload": "plane 256x256, full 3600 s", "precision": "fp64 CPU oracle",
           "steps": r.steps, "seconds": round(r.step_seconds, 4),
           "gcups": round(plane.cells * r.steps / r.step_seconds / 1e9, 4)}
    out["rows"].append(row)
    print(f"[bench] configs {row}", file=sys.stderr, flush=True)
    del r
    valley = generators.valley_case(n=4096)
    for prec, B in (("fp32", 28.0), ("fp64", 56.0)):
        wall, kern, _ = measure(valley, prec, args.steps, args.warmup, args.tile_rows, args.wpb)
        row = {"workload": f"valley 4096x4096, steps {args.warmup + 1}-{args.warmup + args.steps}",
               "precision": prec, "steps": args.steps, "ms_per_step": round(wall * 1e3 / args.steps, 4),
               "gcups": round(valley.cells * args.steps / wall / 1e9, 3),
               "kernel_ms": round(kern * 1e3, 4),
               "roofline_frac": round(B * valley.cells / kern / 1e9 / HBM_PEAK_GBS, 4)}
        out["rows"].append(row)
        print(f"[bench] configs {row}", file=sys.stderr, flush=True)
    r = oracle.OracleRun(valley, dt_log_cap=64, mass_log_cap=4)
    r.advance(10)
    warm = r.step_seconds
    r.advance(30)
    secs = r.step_seconds - warm
    row = {"workload": "valley 4096x4096, steps 11-40", "precision": "fp64 CPU oracle",
           "steps": 30, "seconds": round(secs, 3),
           "gcups": round(valley.cells * 30 / secs / 1e9, 4)}
    out["rows"].append(row)
    print(f"[bench] configs {row}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def bench_shard(args):
    """One GPU's share of the 16384^2 roofline case split into `--parts` row
    slabs: slab 0 (16384 x 2048 at 8 parts) with its own point sources.  The
    slab's north and south edges stand in for the halo rows a sharded run
    would exchange; the step's work per cell is the same."""
    from adaptflood_amd import generators
    from adaptflood_amd.scenario import Scenario
    full = generators.roofline_case(n=args.size)
    ny = full.ny // args.parts
    pts = [p for p in full.points if p.y < ny]
    slab = Scenario(dem=full.dem[:ny].copy(), dx=full.dx, sim_time=full.sim_time,
                    west_free=True, east_free=True, north_free=True, south_free=False,
                    points=pts, name=f"slab {full.nx}x{ny}")
    del full
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    cells = slab.cells
    res = {}
    for prec, B in (("fp32", 28.0), ("fp64", 56.0)):
        wall, kern, _ = measure(slab, prec, args.steps, args.warmup, args.tile_rows, args.wpb)
        res[prec] = {"value": round(cells * args.steps / wall / 1e9, 3), "unit": "GCUPS",
                     "ms_per_step": round(wall * 1e3 / args.steps, 4),
                     "roofline": {"bound": "hbm", "achieved": round(B * cells / kern / 1e9, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(B * cells / kern / 1e9 / HBM_PEAK_GBS, 4),
                                  "kernel_ms": round(kern * 1e3, 4)}}
        print(f"[bench] shard {prec}: {res[prec]}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "GCUPS of one GPU's row slab", "config": {
        "workload": f"row slab {slab.nx}x{slab.ny} (1/{args.parts} of the {args.size}^2 "
                    f"roofline case), {len(pts)} QFIX", "cells": cells}, **res}), flush=True)


def bench_ceiling(args):
    """The HBM bandwidth the step's memory pattern reaches with no stencil and
    no arithmetic (af_hbm_probe, built from adaptflood_amd/csrc/hbm_probe.hip),
    beside a read-only stream and a copy: the ceiling the roofline fraction of
    the step is measured against."""
    import subprocess
    exe = os.path.join(ROOT, "adaptflood_amd", "_lib", "af_hbm_probe")
    r = subprocess.run([exe, str(args.size), "64", "2"], capture_output=True, text=True,
                       timeout=300)
    sys.stderr.write(r.stdout)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        sys.exit(r.returncode)
    best = {}
    for line in r.stdout.splitlines():
        f = line.split()
        name, ms, gbs = f[0], float(f[1]), float(f[3])
        if name not in best or gbs > best[name]["GB/s"]:
            best[name] = {"ms": ms, "GB/s": gbs, "frac_of_8TB/s": round(gbs / HBM_PEAK_GBS, 4)}
    print(json.dumps({"metric": "HBM bandwidth of the step's access pattern without the step",
                      "grid": f"{args.size}x{args.size}, 256-column strips, 64-row tiles",
                      "patterns": best}), flush=True)


def measure_wet(scn, args):
    """FP64 step time on the roofline workload once it has wetted: the domain
    runs to --wet-time simulated seconds, then `steps` steps are timed (the
    FP64 step evaluates the exact pow(h, 10/3) only in wavefronts with a wet
    face, so it costs more as the flood spreads)."""
    import torch
    from adaptflood_amd.domain import Domain, make_tuning
    d = Domain(scn, precision="fp64", tuning=make_tuning(graph_steps=216))
    d.set_sim_time(args.wet_time)
    d.run(check_every=2160, graph=True)
    d.set_sim_time(scn.sim_time)
    H = d.read("H")
    wet = float(np.count_nonzero(H > 0.01)) / H.size
    d.advance(0, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.advance(args.steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = _event_ms(torch, d.launch_step, min(args.steps, 50)) * 1e-3
    info = d.info()
    variant = d.variant()
    d.close()
    cells = scn.cells
    a64 = 56.0 * cells / kern / 1e9
    print(f"[bench] fp64 wet (t={info['t']:.0f}s, {wet * 100:.1f}% of cells wet): "
          f"{wall * 1e3 / args.steps:.3f} ms/step, {kern * 1e3:.3f} ms/launch, {a64:.0f} GB/s",
          file=sys.stderr, flush=True)
    return {"value": round(cells * args.steps / wall / 1e9, 3), "unit": "GCUPS",
            "ms_per_step": round(wall * 1e3 / args.steps, 4),
            "state": f"after {args.wet_time:.0f} s of the run, {wet * 100:.1f}% of cells wet "
                     f"(H > 0.01 m)",
            "variant": variant,
            "roofline": {"bound": "hbm", "achieved": round(a64, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(a64 / HBM_PEAK_GBS, 4),
                         "kernel_ms": round(kern * 1e3, 4)}}


def bench_batched_leg(scns, args):
    """The 512-scenario ensemble in one launch per step (autotuned tiling)
    against every 4th scenario run alone through the fused single-domain path."""
    import torch
    from adaptflood_amd.domain import Batch, make_tuning
    from adaptflood_amd.tune import autotune
    n, cells = len(scns), sum(s.cells for s in scns)
    t = time.perf_counter()
    best, _ = autotune(scns, "fp32", rows=(8, 16, 32, 64), launches=10)
    tuning_s = time.perf_counter() - t
    tun = make_tuning(tile_rows=best.tile_rows, waves_per_block=best.waves_per_block,
                      variant=best.variant, graph_steps=36, tile_order=best.tile_order)
    b = Batch(scns, precision="fp32", tuning=tun)
    b.advance(args.warmup, graph=False)
    b.advance(0, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(args.steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.close()
    del b
    torch.cuda.empty_cache()
    gcups = cells * args.steps / wall / 1e9
    sample = scns[::args.seq_stride]
    g_seq, _ = sequential_fused(sample, "fp32", args.seq_steps)
    print(f"[bench] batched ensemble: {gcups:.1f} GCUPS ({wall * 1e3 / args.steps:.3f} ms/step,"
          f" rows {best.tile_rows} variant {best.variant}); sequential fused {g_seq:.2f} GCUPS",
          file=sys.stderr, flush=True)
    return {"value": round(gcups, 3), "unit": "GCUPS",
            "ms_per_step": round(wall * 1e3 / args.steps, 4),
            "workload": f"{n} ragged scenarios from the ensemble generator (seed 4000, sides "
                        f"log-uniform 128..2048, aspect U(1,4), {cells} cells), FP32, one launch "
                        f"per step, {args.steps} steps after {args.warmup}",
            "tiling": {"tile_rows": best.tile_rows, "variant": best.variant,
                       "tile_order": best.tile_order, "autotune_s": round(tuning_s, 1)},
            "sequential_fused": {"value": round(g_seq, 3), "unit": "GCUPS",
                                 "sample": f"every {args.seq_stride}th scenario ({len(sample)}), "
                                           f"{args.seq_steps} steps each, alone"},
            "speedup_vs_sequential_fused": round(gcups / g_seq, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0,
                    help="served: the pool's kernel variant (0: the library's default)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--only-fp64", action="store_true",
                    help="measure only the FP64 step (for profiling runs)")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="CPU oracle steps of the baseline sample, after 10 untimed ones "
                         "(about 18 s on 16 threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="leave out the batched 512-scenario ensemble key")
    ap.add_argument("--wet-time", type=float, default=1800.0,
                    help="fp64_wet: simulated seconds run before timing FP64 (0: leave out)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the reference-call-sequence timing of the default workload")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served-ensemble latency run of the default workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 4 GiB device-to-device copy (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--workload",
                    choices=("roofline", "ensemble", "served", "accuracy", "divergence", "shard",
                             "driver", "lone", "twin", "ceiling", "configs", "budget"),
                    default="roofline")
    ap.add_argument("--parts", type=int, default=8, help="shard: row slabs of the roofline case")
    ap.add_argument("--twin-cases", default="roofline:2048:3600,valley:4096:7200,"
                    "roofline:16384:3600", help="twin: case:size:sim_time list")
    ap.add_argument("--twin-nudge", choices=("source", "dem"), default="dem",
                    help="twin: one source value +1 ulp, or every DEM value +1 ulp")
    ap.add_argument("--budget-size", type=int, default=2048,
                    help="budget: side of the roofline-generator grid (CPU oracle)")
    ap.add_argument("--budget-steps", type=int, default=1600,
                    help="budget: steps of each oracle run")
    ap.add_argument("--budget-qfix", type=int, default=0,
                    help="budget: QFIX sources (0: the generator's 5000; 78 keeps the 16384^2 "
                         "case's sources per cell at 2048^2)")
    ap.add_argument("--twin-linear", type=int, default=3200,
                    help="twin: checkpoint spacing once the doubling schedule reaches it")
    ap.add_argument("--valley-size", type=int, default=4096, help="accuracy: valley case side")
    ap.add_argument("--acc-time", type=float, default=600.0,
                    help="accuracy: simulated seconds of the valley comparison with the oracle")
    ap.add_argument("--roof-time", type=float, default=3600.0,
                    help="accuracy: simulated seconds of the roofline-case FP32/FP64 comparison")
    ap.add_argument("--drv-time", type=float, default=7200.0,
                    help="driver: simulated seconds of the valley case")
    ap.add_argument("--save-int", type=float, default=600.0, help="driver: seconds between saves")
    ap.add_argument("--div-cases", default="valley,roofline",
                    help="divergence: comma-separated cases to run (valley, roofline)")
    ap.add_argument("--acc-cases", default="valley,roofline",
                    help="accuracy: comma-separated cases to run (valley, roofline)")
    ap.add_argument("--members", type=int, default=512, help="ensemble / served scenarios")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32",
                    help="ensemble / served state precision")
    ap.add_argument("--seq-stride", type=int, default=4,
                    help="sequential fused baseline: every n-th scenario")
    ap.add_argument("--seq-steps", type=int, default=72)
    ap.add_argument("--plain-stride", type=int, default=16,
                    help="sequential plain baseline: every n-th scenario")
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--lone-steps", type=int, default=720, help="lone: timed steps per domain")
    ap.add_argument("--rate", type=float, default=20.0, help="served arrivals per second")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="served HBM budget (GiB)")
    ap.add_argument("--chunk", type=int, default=36, help="served steps per pool advance")
    ap.add_argument("--lane", action="store_true",
                    help="served: move requests that need many more steps than the rest to the "
                         "pool's lane (the default bench's saturated leg always does)")
    ap.add_argument("--autotune", action="store_true",
                    help="ensemble: pick tile rows and kernel variant by timing candidates")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "accuracy":
        return bench_accuracy(args)
    if args.workload == "divergence":
        return bench_divergence(args)
    if args.workload == "shard":
        return bench_shard(args)
    if args.workload == "driver":
        return bench_driver(args)
    if args.workload == "lone":
        return bench_lone(args)
    if args.workload == "twin":
        return bench_twin(args)
    if args.workload == "configs":
        return bench_configs(args)
    if args.workload == "budget":
        return bench_budget(args)
    if args.workload == "ceiling":
        return bench_ceiling(args)
    batched_scns = None
    if not args.no_batched:
        t = time.perf_counter()
        batched_scns = generate_members(list(range(args.members)), 4000, [3600.0] * args.members)
        print(f"[bench] generated {len(batched_scns)} ensemble scenarios in "
              f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    served_reqs = None
    if not args.no_served:
        # generated in worker processes before this process touches the GPU
        from adaptflood_amd import generators as gen
        arr, sims = gen.arrivals(args.members, rate=args.rate, seed=5)
        t = time.perf_counter()
        reqs = generate_members(list(range(args.members)), 5000, [float(x) for x in sims])
        served_reqs = (reqs, arr)
        print(f"[bench] generated {len(reqs)} served requests in {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
    import torch
    from adaptflood_amd import generators
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    t = time.perf_counter()
    scn = generators.roofline_case(n=args.size)
    print(f"[bench] generated {scn.nx}x{scn.ny} in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    cells = scn.cells
    if args.calibrate_copy:
        a = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    if args.only_fp64:
        wall64, kern64, _ = measure(scn, "fp64", args.steps, args.warmup, args.tile_rows, args.wpb)
        print(f"[bench] fp64: {wall64 * 1e3 / args.steps:.3f} ms/step wall, "
              f"{kern64 * 1e3:.3f} ms/launch, {56.0 * cells / kern64 / 1e9:.0f} GB/s",
              file=sys.stderr, flush=True)
        return
    wall, kern_s, info = measure(scn, "fp32", args.steps, args.warmup, args.tile_rows, args.wpb)
    gcups = cells * args.steps / wall / 1e9
    achieved = 28.0 * cells / kern_s / 1e9
    print(f"[bench] fp32: {wall * 1e3 / args.steps:.3f} ms/step wall, {kern_s * 1e3:.3f} ms/launch,"
          f" {gcups:.1f} GCUPS, {achieved:.0f} GB/s, t={info['t']:.1f}s steps={info['steps']}",
          file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"roofline {scn.nx}x{scn.ny} at 2 m, fractal DEM, 10% NODATA, "
                               f"5000 QFIX, FREE edges, FP32 state, dry start",
                   "cells": cells, "precision": "fp32"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": _traffic_from_profile("fp32"),
                     "algorithmic_bytes_per_launch": 28 * cells,
                     "kernel_ms": round(kern_s * 1e3, 4)},
    }
    if not args.no_fp64:
        wall64, kern64, info64 = measure(scn, "fp64", args.steps, args.warmup, args.tile_rows,
                                         args.wpb)
        g64 = cells * args.steps / wall64 / 1e9
        a64 = 56.0 * cells / kern64 / 1e9
        print(f"[bench] fp64: {wall64 * 1e3 / args.steps:.3f} ms/step wall, "
              f"{kern64 * 1e3:.3f} ms/launch, {g64:.1f} GCUPS, {a64:.0f} GB/s",
              file=sys.stderr, flush=True)
        out["fp64"] = {"value": round(g64, 3), "unit": "GCUPS",
                       "ms_per_step": round(wall64 * 1e3 / args.steps, 4),
                       "roofline": {"bound": "hbm", "achieved": round(a64, 1),
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": round(a64 / HBM_PEAK_GBS, 4),
                                    "traffic": _traffic_from_profile("fp64"),
                                    "kernel_ms": round(kern64 * 1e3, 4)}}
    if not args.no_fp64 and args.wet_time > 0:
        out["fp64_wet"] = measure_wet(scn, args)
    if batched_scns is not None:
        out["batched"] = bench_batched_leg(batched_scns, args)
    if not args.no_plain:
        # the reference's own call sequence (per-function FP64 kernels, dt read
        # back every step) on the same grid, for scale
        from adaptflood_amd import kernels
        r = kernels.run_sequence(scn, max_steps=args.plain_steps)
        g = cells * r["steps"] / r["loop_seconds"] / 1e9
        out["plain_path"] = {"value": round(g, 3), "unit": "GCUPS",
                             "ms_per_step": round(r["loop_seconds"] * 1e3 / r["steps"], 3),
                             "sample": f"reference call sequence through the per-function C "
                                       f"ABI, FP64, first {r['steps']} steps of the same workload"}
        print(f"[bench] plain path fp64: {g:.2f} GCUPS, "
              f"{r['loop_seconds'] * 1e3 / r['steps']:.2f} ms/step", file=sys.stderr, flush=True)
        del r
        torch.cuda.empty_cache()
    if served_reqs is not None:
        reqs, arr = served_reqs
        done, summary, budget = run_served(reqs, arr, "fp32", args.budget_gb, args.chunk)
        out["served"] = {
            "workload": f"{len(reqs)} requests from the ensemble generator (seed 5), Poisson "
                        f"arrivals at {args.rate}/s, sim_time U(600,3600) s, FP32, HBM budget "
                        f"{budget / (1 << 30):.0f} GiB, continuous batching",
            "p50_latency_s": round(summary["p50_latency"], 3),
            "p99_latency_s": round(summary["p99_latency"], 3),
            "gcups": round(summary["gcups"], 3), "wall_s": round(summary["wall_s"], 3),
            "arrival_span_s": round(float(arr[-1]), 3)}
        # the same requests arriving at 1000/s: the pool saturates and its
        # aggregate rate, not the arrivals, sets the wall time (the arrival
        # draw leaves the per-request sim_times unchanged)
        from adaptflood_amd import generators as gen
        arr_sat, _ = gen.arrivals(len(reqs), rate=1000.0, seed=5)
        done, summary, budget = run_served(reqs, arr_sat, "fp32", args.budget_gb, args.chunk,
                                           lane=True)
        # the same requests one after another through the single-scenario
        # fused path: every seq_stride-th request, timed alone over
        # seq_steps steps, aggregated uniformly and by its served step count
        own = {r.rid: r.steps for r in done}
        sample = [k for k in range(0, len(reqs), args.seq_stride) if k in own]
        g_fused, g_own, s_own = sequential_fused_own(
            [reqs[k] for k in sample], [own[k] for k in sample], "fp32", args.seq_steps)
        print(f"[bench] sequential fused fp32 ({len(sample)} requests): {g_fused:.2f} GCUPS "
              f"({args.seq_steps} steps each), {g_own:.2f} GCUPS (own step counts)",
              file=sys.stderr, flush=True)
        out["served_saturated"] = {
            "workload": f"the same {len(reqs)} requests, Poisson arrivals at 1000/s, with the "
                        f"pool's lane",
            "p50_latency_s": round(summary["p50_latency"], 3),
            "p99_latency_s": round(summary["p99_latency"], 3),
            "gcups": round(summary["gcups"], 3), "wall_s": round(summary["wall_s"], 3),
            "arrival_span_s": round(float(arr_sat[-1]), 3),
            "sequential_fused": {
                "value": round(g_fused, 3), "unit": "GCUPS",
                "sample": f"every {args.seq_stride}th request ({len(sample)}), alone, "
                          f"{args.seq_steps} graph-replayed steps each"},
            "sequential_fused_own_steps": {
                "value": round(g_own, 3), "unit": "GCUPS",
                "sample": "the same requests at their measured lone step times, each for the "
                          "steps it took when served (run to completion one after another)",
                "seconds": round(s_own, 3)},
            "speedup_vs_sequential_fused": round(summary["gcups"] / g_fused, 2),
            "speedup_vs_sequential_fused_own_steps": round(summary["gcups"] / g_own, 2)}
    if not args.no_cpu_baseline:
        from oracle import oracle
        r = oracle.OracleRun(scn, dt_log_cap=args.cpu_steps + 11, mass_log_cap=4)
        r.advance(10)                      # untimed warm-up steps
        warm_s = r.step_seconds
        r.advance(args.cpu_steps)
        secs = r.step_seconds - warm_s
        cpu_gcups = cells * args.cpu_steps / secs / 1e9
        host = host_cpu()
        out["cpu_baseline"] = {
            "value": round(cpu_gcups, 4), "unit": "GCUPS", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"FP64 OpenMP oracle, steps 11-{10 + args.cpu_steps} of the same "
                      f"{scn.nx}x{scn.ny} workload after 10 untimed steps ({secs:.1f} s)",
            "host_cpu": host["model"], "host_physical_cores": host["physical_cores"],
            "host_nproc": host["nproc"]}
        print(f"[bench] cpu: {cpu_gcups:.3f} GCUPS on {oracle.num_threads()} threads",
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
