"""Benchmark: GCUPS and HBM-roofline fraction of the fused local-inertial step.

Default workload (one MI355X): the 16384 x 16384 roofline case at 2 m cells,
fractal DEM with 10% NODATA blobs, 5000 QFIX sources, all edges FREE, FP32
state, starting dry; W untimed warm-up steps, then K timed steps.  Prints one
JSON line.  The same K steps are also measured at FP64 (key "fp64"), and the
FP64 CPU oracle is timed on a bounded sample of the same workload
("cpu_baseline").

Algorithmic bytes per cell-update: read z, H, Qx_old, Qy_old and write H, Qx,
Qy = 7 values = 28 B at FP32 (56 B at FP64).
"""
import argparse
import json
import os
import sys
import time

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--only-fp64", action="store_true",
                    help="measure only the FP64 step (for profiling runs)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 4 GiB device-to-device copy (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

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
    if not args.no_cpu_baseline:
        from oracle import oracle
        r = oracle.OracleRun(scn, dt_log_cap=args.cpu_steps + 1, mass_log_cap=4)
        r.advance(args.cpu_steps)
        cpu_gcups = cells * r.steps / r.step_seconds / 1e9
        out["cpu_baseline"] = {
            "value": round(cpu_gcups, 4), "unit": "GCUPS", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"FP64 OpenMP oracle, first {r.steps} steps of the same "
                      f"{scn.nx}x{scn.ny} workload ({r.step_seconds:.1f} s)"}
        print(f"[bench] cpu: {cpu_gcups:.3f} GCUPS on {oracle.num_threads()} threads",
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
