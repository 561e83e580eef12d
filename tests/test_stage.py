"""Admission staging (af_pool_stage) on the CPU: it needs neither a pool nor a
GPU, so a server can run it on worker threads while the pool steps.  A valid
scenario stages in either precision; one the pool would refuse (a non-finite
DEM value or initial depth, a point off the grid) raises RequestError with the
library's message, from any thread."""
import concurrent.futures
import dataclasses

import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd._lib import AF_FP32, AF_FP64
from adaptflood_amd.scenario import PointBC
from adaptflood_amd.serve import RequestError, Staged


@pytest.mark.parametrize("prec", [AF_FP32, AF_FP64])
def test_stage_valid_scenario(prec):
    s = generators.small_case(70, 33, seed=5)
    st = Staged(s, prec)
    assert st.handle and st.scenario is s
    st.free()
    assert st.handle is None
    st.free()                                   # a second free is a no-op


def test_stage_refuses_bad_inputs():
    s = generators.small_case(40, 30, seed=6)
    dem = np.array(s.dem, dtype=np.float64)
    dem[4, 7] = np.nan
    with pytest.raises(RequestError, match="DEM value at column 7, row 4 is not finite"):
        Staged(dataclasses.replace(s, dem=dem), AF_FP32)
    h0 = np.zeros_like(dem)
    h0[2, 3] = np.inf
    with pytest.raises(RequestError, match="initial depth at column 3, row 2"):
        Staged(dataclasses.replace(s, h0=h0), AF_FP64)
    off = dataclasses.replace(s, points=list(s.points) + [PointBC(0, 30, 1, 0.1)])
    with pytest.raises(RequestError, match="outside the 40 x 30 grid"):
        Staged(off, AF_FP32)


def test_stage_on_worker_threads():
    scns = [generators.small_case(50 + k, 20 + k, seed=10 + k) for k in range(8)]
    bad = np.array(scns[3].dem, dtype=np.float64)
    bad[0, 0] = np.inf
    scns[3] = dataclasses.replace(scns[3], dem=bad)
    with concurrent.futures.ThreadPoolExecutor(4) as ex:
        futs = [ex.submit(Staged, s, AF_FP32) for s in scns]
        for k, f in enumerate(futs):
            if k == 3:
                with pytest.raises(RequestError, match="not finite"):
                    f.result()
            else:
                assert f.result().handle
