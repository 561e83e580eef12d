"""The command-line driver (lisflood_amd) without a GPU: argument handling
and input errors end cleanly before any device call."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adaptflood_amd",
                   "_lib", "lisflood_amd")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.fail("lisflood_amd is not built (run __graft_entry__.build())")
    return EXE


def test_usage(exe, tmp_path):
    r = subprocess.run([exe, "--path", "fp16", "x.par"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    # no .par file: the defaults, inputs from the working directory (main.cpp:28-31);
    # an empty directory has no dem.ascii, which ends the run before any GPU work
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60, cwd=str(tmp_path))
    assert "Using default parameters..." in r.stdout
    assert "  DEM file:    dem.ascii" in r.stdout and "  Sim time:    3600 s" in r.stdout
    assert r.returncode == 1 and "lisflood_amd: load" in r.stderr


def test_missing_par_file(exe, tmp_path):
    r = subprocess.run([exe, str(tmp_path / "none.par")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "lisflood_amd: load" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before any
    GPU work."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["python", os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
