"""Writes the digests of the FP64 oracle's full run of one workload (--case):
the 4096 x 4096 valley workload (fractal terrain with a carved valley at
5 m, 2000 QVAR hydrographs, 20 mm/h rain over 0-1800 s, HFIX 2 m west edge,
FREE N/E/S, 7200 s; the default), the roofline generator at 1024 x 1024, or
a member of the ensemble generator (CASES below).

The valley's full run is 76,456 steps, hours of the CPU oracle (about
0.09 s per step on 16 cores), so a fixture keeps SHA-256 digests of the
state instead of the state itself:
H (ny*nx), Qx and Qy ((ny+1)*(nx+1), pitch nx+1) as little-endian float64,
at every CHECK steps and at the end, the digest of the whole dt log, the step
count, the final t and every mass report (t, dt, Vol, Area); its
"oracle_seconds" is the stepping time of the run's last leg.  A GPU test
advances the FP64 fused path by the same step counts and compares digests;
the reports are compared as values (the volume sum's order differs).

Run from the repository root:
    python tests/golden/make_full_run_digests.py [--case C] [--threads N] [--checkpoint DIR]
--case roofline1024 runs the roofline generator at 1024 x 1024 (fractal
terrain at 2 m, 10% NODATA, 5000 QFIX, free edges) for its full 3600 s
instead, into tests/golden/roofline1024_fp64.json; --case memberK one
member of the ensemble generator (K in MEMBERS), into memberK_fp64.json.
With --checkpoint, the oracle state is saved to DIR after every CHECK steps
and a later run with the same DIR resumes from it.  With --emit, the fixture
is written from the checkpoints in DIR so far (no stepping): its "final" is
null until the run has ended, and "dt_sha256" covers the first "dt_steps"
entries of the dt log.

To continue a run on another machine: --pack DIR PACKDIR writes the latest
checkpoint of DIR in a compact form (H, Qx and Qy byte-shuffled and
deflated, about half the size; Qxold and Qyold are Qx and Qy times 1/dx,
as UpdateQs_GPU stores them, and are rebuilt); on the other machine
--from-pack PACKDIR --checkpoint DIR unpacks it into DIR and resumes.
--seconds S stops stepping after about S seconds (recording the state it
stopped at as a checkpoint) and --out FILE writes the (partial) fixture to
FILE after every checkpoint.

--from-fixture FILE resumes from the last checkpoint of a partial fixture
without a saved state (the state of a run is too large to carry between
machines): the FP64 fused path (GPU) is advanced to that step, and its H, Qx,
Qy and dt log are taken as the oracle's state only if their SHA-256 digests
equal the oracle's own digests of that step; the mass reports and the clock
come from the fixture.  Every later digest is the oracle's own stepping.
"""
import argparse
import hashlib
import json
import os
import shutil
import sys
import time
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from adaptflood_amd import generators  # noqa: E402
from oracle import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
# --case: the workload and its fixture (tests/golden/<case>_fp64.json)
CASES = {
    "valley4096": (lambda: generators.valley_case(n=4096),
                   "valley 4096x4096 at 5 m, 2000 QVAR, rain 0-1800 s, HFIX west, 7200 s"),
    "roofline1024": (lambda: generators.roofline_case(n=1024),
                     "roofline generator 1024x1024 at 2 m, fractal DEM, 10% NODATA, 5000 QFIX, "
                     "FREE edges, 3600 s"),
}
# members of the 512-scenario ensemble generator (seed 4000 + k), run alone
# for their full 3600 s; a GPU test steps them together in one batch
MEMBERS = (3, 21, 31, 48, 69, 78, 151, 241, 292, 354, 374)
for _k in MEMBERS:
    CASES[f"member{_k}"] = (lambda k=_k: generators.ensemble_member(k),
                            f"ensemble generator member {_k} (seed {4000 + _k}), 3600 s")
CASE = "valley4096"
OUT = os.path.join(HERE, f"{CASE}_fp64.json")
CHECK = 5000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def save_state(d, r, checkpoints):
    os.makedirs(d, exist_ok=True)
    tmp = os.path.join(d, "state.tmp.npz")
    np.savez(tmp, H=r.H, Qx=r.Qx, Qy=r.Qy, Qxold=r.Qxold, Qyold=r.Qyold,
             scalars=np.array([r.t, r.next_mass, r.steps, r.step_seconds]),
             dts=np.array(r.dts), reports=np.array(r.reports).reshape(-1, 4))
    os.replace(tmp, os.path.join(d, "state.npz"))
    with open(os.path.join(d, "checkpoints.json"), "w") as f:
        json.dump(checkpoints, f)


def load_state(d, r):
    p = os.path.join(d, "state.npz")
    if not os.path.exists(p):
        return []
    with np.load(p, allow_pickle=False) as z:
        for k in ("H", "Qx", "Qy", "Qxold", "Qyold"):
            getattr(r, k)[:] = z[k]
        t, nm, steps, secs = z["scalars"].tolist()
        r.t, r.next_mass, r.steps, r.step_seconds = t, nm, int(steps), secs
        r.dts = z["dts"].tolist()
        r.reports = z["reports"].tolist()
    with open(os.path.join(d, "checkpoints.json")) as f:
        return json.load(f)


def emit(d):
    """The fixture from the checkpoints saved so far."""
    with open(os.path.join(d, "checkpoints.json")) as f:
        checkpoints = json.load(f)
    with np.load(os.path.join(d, "state.npz"), allow_pickle=False) as z:
        dts = z["dts"]
        reports = z["reports"]
        steps = int(z["scalars"][2])
    assert checkpoints and checkpoints[-1]["steps"] == steps == len(dts)
    scn = CASES[CASE][0]()
    write_fixture(scn, checkpoints, None, dts, reports, None)


def write_fixture(scn, checkpoints, final, dts, reports, seconds):
    out = {
        "workload": CASES[CASE][1],
        "dem_sha256": hashlib.sha256(np.ascontiguousarray(scn.dem, dtype=np.float64)
                                     .tobytes()).hexdigest(),
        "check_every": CHECK,
        "checkpoints": checkpoints,
        "final": final,
        "dt_steps": int(len(dts)),
        "dt_sha256": sha(dts),
        "dt_first": np.asarray(dts[:8]).tolist(),
        "dt_last": np.asarray(dts[-8:]).tolist(),
        "reports": np.asarray(reports).tolist(),
        "oracle_seconds": seconds,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    end = final or checkpoints[-1]
    print(f"wrote {OUT}: {end['steps']} steps, t={end['t']:.6f}"
          + ("" if final else " (partial)"))


PACKED = ("H", "Qx", "Qy")


def pack(src, dst):
    """The latest checkpoint of `src` in compact form, in `dst`."""
    os.makedirs(dst, exist_ok=True)
    with np.load(os.path.join(src, "state.npz"), allow_pickle=False) as z:
        for k in PACKED:
            a = np.ascontiguousarray(z[k], dtype="<f8")
            shuffled = a.view(np.uint8).reshape(-1, 8).T.tobytes()
            with open(os.path.join(dst, k + ".zsh"), "wb") as f:
                f.write(zlib.compress(shuffled, 6))
        np.savez(os.path.join(dst, "small.npz"), scalars=z["scalars"], dts=z["dts"],
                 reports=z["reports"], sizes=np.array([z["H"].size, z["Qx"].size]))
    shutil.copy(os.path.join(src, "checkpoints.json"), os.path.join(dst, "checkpoints.json"))


def unpack(src, dst, dx):
    """A packed checkpoint from `src` as the checkpoint of `dst`."""
    os.makedirs(dst, exist_ok=True)
    with np.load(os.path.join(src, "small.npz"), allow_pickle=False) as small:
        sizes = dict(zip(PACKED, (int(small["sizes"][0]), int(small["sizes"][1]),
                                  int(small["sizes"][1]))))
        arr = {}
        for k in PACKED:
            with open(os.path.join(src, k + ".zsh"), "rb") as f:
                raw = np.frombuffer(zlib.decompress(f.read()), dtype=np.uint8)
            arr[k] = np.ascontiguousarray(raw.reshape(8, sizes[k]).T).view("<f8").ravel()
        dxinv = 1.0 / dx    # UpdateQs_GPU: q_old = Q * (1.0 / dx)
        tmp = os.path.join(dst, "state.tmp.npz")
        np.savez(tmp, H=arr["H"], Qx=arr["Qx"], Qy=arr["Qy"], Qxold=arr["Qx"] * dxinv,
                 Qyold=arr["Qy"] * dxinv, scalars=small["scalars"], dts=small["dts"],
                 reports=small["reports"])
    os.replace(tmp, os.path.join(dst, "state.npz"))
    shutil.copy(os.path.join(src, "checkpoints.json"), os.path.join(dst, "checkpoints.json"))


def resume(r, fx, H, Qx, Qy, dts):
    """Sets the oracle run `r` to the last checkpoint of the partial fixture
    `fx`, given a state (H, Qx, Qy, dt log) whose digests must equal that
    checkpoint's; returns the fixture's checkpoints."""
    rec = fx["checkpoints"][-1]
    if fx["final"] is not None or fx["dt_steps"] != rec["steps"] or len(dts) != rec["steps"]:
        raise ValueError("the fixture's last checkpoint is not a resumable partial run")
    for name, a in (("H", H), ("Qx", Qx), ("Qy", Qy)):
        if sha(a) != rec[name]:
            raise ValueError(f"{name} at step {rec['steps']} differs from the oracle's")
    if sha(dts) != fx["dt_sha256"]:
        raise ValueError(f"dt log to step {rec['steps']} differs from the oracle's")
    r.H[:] = np.asarray(H, dtype=np.float64).ravel()
    r.Qx[:] = Qx
    r.Qy[:] = Qy
    dxinv = 1.0 / r.s.dx    # UpdateQs_GPU: q_old = Q * (1.0 / dx)
    r.Qxold[:] = r.Qx * dxinv
    r.Qyold[:] = r.Qy * dxinv
    r.t, r.steps = rec["t"], rec["steps"]
    r.dts = np.asarray(dts, dtype=np.float64).tolist()
    r.reports = [list(map(float, x)) for x in fx["reports"]]
    nm = float(r.s.mass_int)   # one report per crossing: next_mass += mass_int
    for _ in r.reports:
        nm += r.s.mass_int
    r.next_mass = nm
    return list(fx["checkpoints"])


def fused_state(scn, steps):
    """H, Qx, Qy and the dt log of the FP64 fused path after `steps` steps."""
    from adaptflood_amd.domain import Domain, make_tuning
    d = Domain(scn, precision="fp64", tuning=make_tuning(dt_log_cap=steps + 64, report_cap=1024))
    d.advance(steps, graph=True)
    if d.info()["steps"] != steps:
        raise RuntimeError("the fused path stopped early")
    out = d.read("H"), d.read("Qx"), d.read("Qy"), d.dt_log()[:steps]
    d.close()
    return out


def record(r):
    st = r.state()
    return st, {"steps": st["steps"], "t": st["t"], "H": sha(st["H"]), "Qx": sha(st["Qx"]),
                "Qy": sha(st["Qy"]), "H_max": float(st["H"].max()),
                "H_sum": float(st["H"].sum())}


def main():
    global OUT, CASE
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="valley4096", choices=sorted(CASES))
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--emit", action="store_true")
    ap.add_argument("--pack", nargs=2, metavar=("DIR", "PACKDIR"), default=None)
    ap.add_argument("--from-pack", default=None)
    ap.add_argument("--seconds", type=float, default=0.0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--from-fixture", default=None)
    a = ap.parse_args()
    CASE = a.case
    OUT = os.path.abspath(a.out) if a.out else os.path.join(HERE, f"{CASE}_fp64.json")
    if a.pack:
        return pack(*a.pack)
    if a.emit:
        return emit(a.checkpoint)
    if a.threads:
        oracle.set_threads(a.threads)
    scn = CASES[CASE][0]()
    if a.from_pack:
        unpack(a.from_pack, a.checkpoint, scn.dx)
    r = oracle.OracleRun(scn, dt_log_cap=1 << 18, mass_log_cap=1024)
    checkpoints = load_state(a.checkpoint, r) if a.checkpoint else []
    if a.from_fixture:
        with open(a.from_fixture) as f:
            fx = json.load(f)
        steps = fx["checkpoints"][-1]["steps"]
        print(f"fused FP64 path to step {steps}", flush=True)
        checkpoints = resume(r, fx, *fused_state(scn, steps))
        print("its state and dt log equal the oracle's digests", flush=True)
    if checkpoints:
        print(f"resumed at step {r.steps}, t={r.t:.3f}", flush=True)
    t0 = time.time()
    final = None
    # pieces of at most 500 steps, so that --seconds is kept to within a piece
    while r.t < scn.sim_time:
        out_of_time = a.seconds and time.time() - t0 > a.seconds
        if not out_of_time:
            nxt = (r.steps // CHECK + 1) * CHECK
            r.advance(min(500, nxt - r.steps))
            if r.steps < nxt and r.t < scn.sim_time:
                print(f"{time.time() - t0:8.0f} s  steps={r.steps} t={r.t:.3f}", flush=True)
                continue
        elif checkpoints and checkpoints[-1]["steps"] == r.steps:
            break
        st, rec = record(r)
        print(f"{time.time() - t0:8.0f} s  steps={st['steps']} t={st['t']:.3f} "
              f"H_max={rec['H_max']:.6f}", flush=True)
        if r.t < scn.sim_time:
            checkpoints.append(rec)
            if a.checkpoint:
                save_state(a.checkpoint, r, checkpoints)
        else:
            final = rec
        if a.out or final:
            write_fixture(scn, checkpoints, final, st["dt"], st["reports"],
                          time.time() - t0 if final else None)
        if out_of_time:
            break

if __name__ == "__main__":
    main()
