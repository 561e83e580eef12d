"""lisflood_amd end to end on a small case written in the reference formats:
the compat path (the reference's call sequence) and the fused FP64/FP32 paths
against the oracle, through the files and console lines the reference writes."""
import os
import re
import subprocess

import numpy as np
import pytest

from adaptflood_amd import formats, generators
from oracle import oracle

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adaptflood_amd",
                   "_lib", "lisflood_amd")
REPORT = re.compile(r"^t=(\d+)s: Tstep=([\d.]+), Vol=(\d+)m3, Area=(\d+)m2$")


def _read_asc(path):
    with open(path) as f:
        hdr = [f.readline() for _ in range(6)]
        data = np.loadtxt(f)
    return hdr, data


def _save_count(dts, save_int, sim_time):
    """H_%04d files main.cpp:327-363 writes: one at t=0, then one after every
    step that ends at t >= next_save."""
    t, nxt, n = 0.0, save_int, 1
    for dt in dts:
        t += dt
        if t >= nxt:
            n += 1
            nxt += save_int
    return n


def _oracle_stages(scn, cells, interval):
    """Gauge lines the reference writes: after every step that ends at
    t >= next_stage (main.cpp:306-322)."""
    r = oracle.OracleRun(scn)
    nxt, lines = interval, []
    while r.t < scn.sim_time:
        r.advance(1)
        if r.t >= nxt:
            lines.append([f"{r.t:.1f}"] + [f"{r.H[k]:.4f}" if 0 <= k < r.H.size else "NaN"
                                           for k in cells])
            nxt += interval
    return lines


@pytest.mark.parametrize("path,tol", [("compat", 1e-4), ("fp64", 1e-4), ("fp32", 1.1e-3)])
def test_driver_paths(gpu, tmp_path, path, tol):
    scn = generators.small_case(48, 36, seed=17, sim_time=300.0)
    scn.mass_int = 30.0
    for k, p in enumerate(scn.points):
        p.name = f"pt{k}"
    case = tmp_path / "case"
    case.mkdir()
    # gauges at cells (5, 3) and (40, 30), one east of the grid at (60, 1),
    # whose flat index still falls inside (main.cpp's only range test)
    stage = case / "gauges.txt"
    tly = 36 * scn.dx
    stage.write_text("# x y name\n"
                     f"{5.5 * scn.dx} {tly - 3.5 * scn.dx} g1\n"
                     f"{40.5 * scn.dx} {tly - 30.5 * scn.dx}\n"
                     f"{60.5 * scn.dx} {tly - 1.5 * scn.dx} east\n")
    par = formats.write_case(str(case), scn, extra_par=f"saveint 100\nvoutput\n"
                                                     f"stagefile {stage}\nstageint 50\n")
    back = formats.load_case(par).scenario()
    ref = oracle.run(back)
    print(f"driver --path {path}: oracle steps={ref['steps']}", flush=True)
    r = subprocess.run([EXE, "--path", path, "--quiet", par], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    # the console lines of main.cpp:19-46, 70-86, 109, 163, 393
    n_flow = sum(1 for p in back.points if p.type in (oracle.QFIX, oracle.QVAR))
    for line in ("=== LISFLOOD-GPU Flood Simulation ===", "Parameters:", "  Sim time:    300 s",
                 "  Save int:    100 s", "  Velocity:    enabled", "  Async I/O:   enabled",
                 f"Flow boundaries: {n_flow}, Other boundaries: {len(back.points) - n_flow}",
                 "Starting simulation: 300 seconds, output every 100 seconds",
                 "Stage output saved.", "=== Simulation Complete ==="):
        assert line in r.stdout.splitlines(), line
    out = str(tmp_path / "case" / "results")
    _, H = _read_asc(os.path.join(out, "H_final.asc"))
    err = float(np.max(np.abs(H - ref["H"])))
    print(f"driver --path {path}: max|dH| (4-decimal file) = {err:.2e}", flush=True)
    assert err <= tol
    assert os.path.exists(os.path.join(out, "Vx_final.asc"))
    assert os.path.exists(os.path.join(out, "Vy_final.asc"))
    n_saved = len([f for f in os.listdir(out) if re.match(r"H_\d{4}\.asc$", f)])
    reps = [REPORT.match(line) for line in r.stdout.splitlines()]
    reps = [m.groups() for m in reps if m]
    assert len(reps) == len(ref["reports"])
    with open(os.path.join(out, "stage.csv")) as f:
        rows = [line.strip().split(",") for line in f]
    assert rows[0] == ["Time(s)", "g1", "Stage2", "east"]
    want = _oracle_stages(back, [5 + 3 * 48, 40 + 30 * 48, 60 + 1 * 48], 50.0)
    assert len(rows) - 1 == len(want)
    if path == "fp32":
        for got, exp in zip(rows[1:], want):
            assert all(abs(float(a) - float(b)) <= tol for a, b in zip(got[1:], exp[1:]))
        return  # FP32 dt may differ in the last bits; the step count is not pinned
    assert rows[1:] == want
    assert n_saved == _save_count(ref["dt"], 100.0, 300.0)
    steps = int(re.search(r"Steps: (\d+)", r.stdout).group(1))
    assert steps == ref["steps"]
    for got, want in zip(reps, ref["reports"]):
        assert got[0] == f"{want[0]:.0f}" and got[1] == f"{want[1]:.2f}"
        assert abs(int(got[2]) - want[2]) <= 1 and got[3] == f"{want[3]:.0f}"


def _write_small_case(tmp_path, extra):
    scn = generators.small_case(40, 30, seed=23, sim_time=240.0)
    scn.mass_int = 60.0
    case = tmp_path / "case"
    case.mkdir(exist_ok=True)
    return formats.write_case(str(case), scn, xll=1234.5, yll=987.25, extra_par=extra)


def _saves(out, prefix):
    return sorted(f for f in os.listdir(out) if re.match(prefix + r"_\d{4}\.asc$", f))


@pytest.mark.parametrize("path", ["compat", "fp32"])
def test_async_and_sync_saves(gpu, tmp_path, path):
    """async_output 1 (the reference default) writes the saves through the
    writer thread in AsyncOutputManager's format (corners %.6f, the grid's
    NODATA value) with Vx/Vy beside each H; async_output 0 writes write_asc's
    format and H only.  The grids of every save agree to the last digit.
    output_format 2 adds output.nc, whose records hold the same grids;
    output_format 1 writes output.nc and no ASC saves."""
    runs = {}
    for mode in (1, 0):
        sub = tmp_path / f"m{mode}"
        sub.mkdir()
        par = _write_small_case(sub, f"saveint 60\nvoutput\nasync_output {mode}\n")
        r = subprocess.run([EXE, "--path", path, par], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        runs[mode] = (str(sub / "case" / "results"), r.stdout)
    (oa, sa), (os_, ss) = runs[1], runs[0]
    assert "Queued: t=0s (initial state)" in sa and "Async output finished." in sa
    assert "Queued:" not in ss and "Saved:" in ss
    ha, hs = _saves(oa, "H"), _saves(os_, "H")
    assert ha == hs and len(ha) >= 4
    assert _saves(oa, "Vx") == [f.replace("H_", "Vx_") for f in ha]
    assert _saves(oa, "Vy") == [f.replace("H_", "Vy_") for f in ha]
    assert _saves(os_, "Vx") == []
    for f in ha:
        hdr_a, da = _read_asc(os.path.join(oa, f))
        hdr_s, ds = _read_asc(os.path.join(os_, f))
        assert hdr_a[2] == "xllcorner     1234.500000\n" and hdr_a[3] == "yllcorner     987.250000\n"
        assert hdr_a[5] == "NODATA_value  -9999\n"
        assert hdr_s[2] == "xllcorner     1234.5\n" and hdr_s[3] == "yllcorner     987.2\n"
        assert np.array_equal(da, ds), f
    _, vx_last = _read_asc(os.path.join(oa, ha[-1].replace("H_", "Vx_")))
    assert np.count_nonzero(vx_last) > 50
    for f in ("H_final.asc", "Vx_final.asc", "Vy_final.asc"):
        with open(os.path.join(oa, f)) as a, open(os.path.join(os_, f)) as b:
            assert a.read() == b.read(), f
    # output_format 2: output.nc (one float record per save, t = 0 included)
    # beside the ASC saves; output_format 1: output.nc only
    from scipy.io import netcdf_file
    for fmt in (2, 1):
        sub = tmp_path / f"nc{fmt}"
        sub.mkdir()
        par = _write_small_case(sub, f"saveint 60\nvoutput\noutput_format {fmt}\n")
        r = subprocess.run([EXE, "--path", path, par], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        out = str(sub / "case" / "results")
        assert "NetCDF file created:" in r.stdout and "NetCDF output:" in r.stdout
        with netcdf_file(os.path.join(out, "output.nc"), "r", mmap=False) as f:
            times = np.array(f.variables["time"][:])
            Hn = np.array(f.variables["H"][:])
            Vxn = np.array(f.variables["Vx"][:])
        assert len(times) == len(ha) and times[0] == 0.0 and np.all(np.diff(times) > 0)
        if fmt == 1:
            assert _saves(out, "H") == [] and os.path.exists(os.path.join(out, "H_final.asc"))
            continue
        assert _saves(out, "H") == ha
        for k, f in enumerate(ha):          # the ASC grids carry 4 decimals
            _, da = _read_asc(os.path.join(out, f))
            _, va = _read_asc(os.path.join(out, f.replace("H_", "Vx_")))
            assert np.max(np.abs(Hn[k] - da)) <= 5.1e-5, f
            assert np.max(np.abs(Vxn[k] - va)) <= 5.1e-5 + 1e-6 * np.max(np.abs(va)), f


def test_runs_on_defaults_without_a_par_file(gpu, tmp_path):
    """No .par file: the defaults of types.hpp, dem.ascii / bci.bci / bdy.bdy
    from the working directory and results/ as output (main.cpp:28-31); the
    run matches the oracle on the same inputs and defaults."""
    scn = generators.small_case(30, 22, seed=29, sim_time=3600.0)
    scn.fpn, scn.cfl, scn.depth_thresh, scn.mass_int = 0.035, 0.7, 0.001, 60.0
    par = formats.write_case(str(tmp_path / "w"), scn)
    back = formats.load_case(par).scenario()
    for src, dst in (("case.dem.ascii", "dem.ascii"), ("case.bci", "bci.bci"),
                     ("case.bdy", "bdy.bdy")):
        os.replace(str(tmp_path / "w" / src), str(tmp_path / "w" / dst))
    r = subprocess.run([EXE, "--path", "fp64", "--quiet"], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path / "w"))
    assert r.returncode == 0, r.stderr
    assert "Using default parameters..." in r.stdout
    back.rain_times = back.rain_rates = None          # the defaults name no rain file
    back.sim_time = 3600.0
    ref = oracle.run(back)
    _, H = _read_asc(str(tmp_path / "w" / "results" / "H_final.asc"))
    assert int(re.search(r"Steps: (\d+)", r.stdout).group(1)) == ref["steps"]
    assert float(np.max(np.abs(H - ref["H"]))) <= 1e-4


def test_fused_console_matches_reference_sequence(gpu, tmp_path):
    """Without --quiet the fused FP64 path prints the reference's "Step"
    lines (from the device dt log) and mass reports in the same order and
    with the same text as the reference call sequence (compat path)."""
    scn = generators.small_case(40, 31, seed=37, sim_time=700.0)
    scn.mass_int = 45.0
    par = formats.write_case(str(tmp_path / "c"), scn, extra_par="saveint 200\n")
    lines = {}
    for path in ("compat", "fp64"):
        r = subprocess.run([EXE, "--path", path, par], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        lines[path] = [ln for ln in r.stdout.splitlines()
                       if ln.startswith("Step ") or REPORT.match(ln)]
    steps = [ln for ln in lines["compat"] if ln.startswith("Step ")]
    assert len(steps) > 12 and steps[10].startswith("Step 100,")
    assert lines["fp64"] == lines["compat"]


def test_every_step_saves_and_reports(gpu, tmp_path):
    """saveint 0 and massint 0: next_save and next_mass never pass t, so the
    reference saves and reports after every step (main.cpp:291-304,
    326-363).  The fused FP64 path writes the same H files and prints the
    same report lines as the compat path (the reference call sequence)."""
    import shutil
    scn = generators.small_case(30, 20, seed=41, sim_time=60.0)
    scn.mass_int = 0.0
    par = formats.write_case(str(tmp_path / "c"), scn, extra_par="saveint 0\nasync_output 0\n")
    res = str(tmp_path / "c" / "results")
    runs = {}
    for path in ("compat", "fp64"):
        r = subprocess.run([EXE, "--path", path, "--quiet", par], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        reps = [ln for ln in r.stdout.splitlines() if REPORT.match(ln)]
        files = _saves(res, "H")
        grids = {f: _read_asc(os.path.join(res, f))[1] for f in files}
        runs[path] = (reps, files, grids, r.stdout)
        shutil.rmtree(res)
    (rc, fc, gc, oc), (rf, ff, gf, of) = runs["compat"], runs["fp64"]
    steps = int(re.search(r"Steps: (\d+)", oc).group(1))
    print(f"every-step run: {steps} steps, {len(fc)} saves, {len(rc)} reports", flush=True)
    assert steps > 5 and len(rc) == steps and len(fc) == steps + 1
    assert rf == rc and ff == fc
    assert "not recorded" not in of
    for f in fc:
        assert np.array_equal(gf[f], gc[f]), f
