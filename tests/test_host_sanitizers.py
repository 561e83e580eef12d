"""The host-side readers and writers (adaptflood_amd/csrc/io.cpp,
ncwriter.cpp) built with AddressSanitizer and UndefinedBehaviorSanitizer by
g++ (host code only; no GPU code is built here) and driven through the C ABI
on a generated case in the reference's formats, a DEM whose tokens take both
conversion routes and whose last token ends the file, a missing file, and
ASC/NetCDF output with NODATA, -0, NaN and infinite values.  Any sanitizer
report fails the test."""
import os
import shutil
import subprocess

import pytest

from adaptflood_amd import formats, generators

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "adaptflood_amd", "csrc")


@pytest.mark.slow
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_readers_and_writers_are_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize_check")
    build = subprocess.run(
        ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
         "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "host_sanitize_check.cpp"),
         os.path.join(CSRC, "io.cpp"), os.path.join(CSRC, "common.cpp"),
         os.path.join(CSRC, "ncwriter.cpp"), "-o", exe, "-lpthread"],
        capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    par = formats.write_case(str(tmp_path / "case"), generators.small_case(300, 200, seed=3))
    toks = ["1.5", "-0", ".5", "1.", "+7.25", "1e3", "nan", "inf",
            "0." + "0" * 25 + "1", "9007199254740993", "12345678901234567890.5", "-9999"]
    dem = tmp_path / "tok.dem.ascii"
    dem.write_text(f"ncols {len(toks)}\nnrows 1\nxllcorner 0\nyllcorner 0\ncellsize 5\n"
                   f"NODATA_value -9999\n" + " ".join(toks))
    out = tmp_path / "out"
    out.mkdir()
    # verify_asan_link_order=0: the runtime need not come first in the
    # process's library list (the environment may preload others)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, par, str(dem), str(out)], capture_output=True, text=True,
                       timeout=300, env=env)
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-3000:]
    assert "AddressSanitizer" not in text and "runtime error" not in text, text[-3000:]
    assert "all calls returned as expected" in r.stdout
