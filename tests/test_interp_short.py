"""interp_series_short, the fused kernel's series evaluation (all loads of a
series of up to four points issued at once), against interp_series, the
restatement of src/io.cpp:299-311, compiled for the host from the same header:
bitwise equal on random series, including unsorted, repeated and NaN times."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_interp_series_short_equals_reference_loop(tmp_path):
    exe = str(tmp_path / "interp_short_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "adaptflood_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "interp_short_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
