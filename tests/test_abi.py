"""The C ABI library: it loads, exports every entry point include/adaptflood.h
declares, and rejects bad arguments before touching a GPU (CPU-only)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from adaptflood_amd import _lib, generators
from adaptflood_amd._lib import AFError, call

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adaptflood.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(af_[A-Za-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (af_[A-Za-z0-9_]+)", out))
    assert set(names) <= exported
    # the Python binding covers the whole header
    assert set(names) == {s[0] for s in _lib.SIGNATURES}


def test_version_and_error_text():
    assert b"gfx950" in _lib.lib().af_version()


def test_reference_surface_rejects_bad_sizes_without_gpu():
    L = _lib.lib()
    rc = L.af_UpdateQs_GPU(None, None, None, None, 1, 5, 1.0, None)
    assert rc == -1 and b"2 x 2" in L.af_last_error()
    ts = C.c_double()
    rc = L.af_FloodplainQ_GPU(None, None, None, None, None, None, 4, 4, 1.0, 1.0, 0.035, 0.001,
                              10.0, 0.01, C.byref(ts), 1.0, 1.0, 0.7, 9.81, 10.0, 1.0, -9999.0,
                              None, None)
    assert rc == -1   # null pointers
    assert L.af_AddFlowBCs_GPU(None, None, None, 0, None) == 0   # nothing to do
    assert L.af_ApplyPointBCs_GPU(None, None, None, None, -1, 4, 1.0, 1.0, 1.0, None) == -1


def _bytes(scn, precision=0, tuning=None):
    pk = scn.pack()
    n = C.c_size_t()
    t = tuning or _lib.Tuning()
    call("af_domain_arena_bytes", C.byref(pk.desc), precision, C.byref(t), C.byref(n))
    return n.value


def test_arena_size_and_validation():
    scn = generators.small_case(40, 30, seed=1)
    b32 = _bytes(scn, 0)
    b64 = _bytes(scn, 1)
    # 7 state arrays of ny*P (+1 row for Qy) at 4 / 8 bytes, P = round_up(nx+1, 64)
    assert b32 >= 7 * 30 * 64 * 4 and b64 >= 7 * 30 * 64 * 8 and b64 > b32
    bad = generators.small_case(40, 30, seed=1)
    bad.theta = 0.7
    with pytest.raises(AFError, match="theta"):
        _bytes(bad)
    bad = generators.small_case(40, 30, seed=1)
    bad.points[0].y = 30  # idx = x + 30*nx is past the last cell
    with pytest.raises(AFError, match="outside"):
        _bytes(bad)
    bad = generators.small_case(40, 30, seed=1)
    bad.points[0].x, bad.points[0].y = 40, 3   # idx = 40 + 3*40: the reference's wrap to (0, 4)
    assert _bytes(bad) > 0
    t = _lib.Tuning()
    t.graph_steps = 7
    with pytest.raises(AFError, match="multiple of 6"):
        _bytes(scn, 0, t)
    t = _lib.Tuning()
    t.waves_per_block = 3
    with pytest.raises(AFError):
        _bytes(scn, 0, t)


def test_create_rejects_small_arena_before_gpu_work():
    scn = generators.small_case(40, 30, seed=1)
    pk = scn.pack()
    h = C.c_void_p()
    buf = (C.c_char * 512)()
    addr = (C.addressof(buf) + 255) // 256 * 256
    rc = _lib.lib().af_domain_create(C.byref(pk.desc), 0, None, C.c_void_p(addr), 256, None,
                                     C.byref(h))
    assert rc == -4 and b"arena too small" in _lib.lib().af_last_error()


def test_non_finite_inputs_are_refused_before_gpu_work():
    """A NaN or infinite parameter, point value, series entry or rain entry
    would reach every cell on the first step (the reference runs it and
    returns NaN); the fused path refuses it when the scenario is sized."""
    import dataclasses
    base = generators.small_case(40, 30, seed=1)
    assert _bytes(base) > 0
    for field in ("fpn", "cfl", "depth_thresh", "init_tstep", "sim_time", "mass_int", "dx"):
        for v in (np.nan, np.inf):
            with pytest.raises(AFError, match="finite"):
                _bytes(dataclasses.replace(base, **{field: v}))
    with pytest.raises(AFError, match="NaN"):
        _bytes(dataclasses.replace(base, nodata=np.nan))
    qfix = next(k for k, p in enumerate(base.points) if p.type == _lib.AF_BC_QFIX)
    qvar = next(k for k, p in enumerate(base.points) if p.type == _lib.AF_BC_QVAR)
    free = next(k for k, p in enumerate(base.points) if p.type == _lib.AF_BC_FREE)
    pts = [dataclasses.replace(p) for p in base.points]
    pts[qfix].value = np.inf
    with pytest.raises(AFError, match="point .*not finite"):
        _bytes(dataclasses.replace(base, points=pts))
    pts = [dataclasses.replace(p) for p in base.points]
    pts[qvar].values = np.array([0.0, np.nan, 0.0])
    with pytest.raises(AFError, match="series entry 1 is not finite"):
        _bytes(dataclasses.replace(base, points=pts))
    pts = [dataclasses.replace(p) for p in base.points]
    pts[free].value = np.nan       # a FREE point's value is never used
    assert _bytes(dataclasses.replace(base, points=pts)) > 0
    with pytest.raises(AFError, match="rain entry 1"):
        _bytes(dataclasses.replace(base, rain_rates=np.array([1e-5, np.inf, 0.0])))
