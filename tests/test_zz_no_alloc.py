"""af_FloodplainQ_GPU_noalloc (H-max key in the library's static device
scratch) against af_FloodplainQ_GPU (key in a caller workspace), through the
C ABI: bitwise equal Tstep, Qx and Qy on two calls in a row, the second on a
smaller grid whose H maximum is lower than the first's, so a key left over
from the first call would give the second a wrong Tstep."""
import numpy as np
import pytest

from adaptflood_amd import kernels as K

pytestmark = pytest.mark.gpu


def _case(seed, nx, ny, hmax):
    rng = np.random.default_rng(seed)
    dem = rng.uniform(0, 3, nx * ny)
    dem[rng.random(nx * ny) < 0.08] = -9999.0
    H = np.maximum(rng.uniform(-0.2, 1.0, nx * ny), 0.0) * hmax
    nq = (nx + 1) * (ny + 1)
    return dem, H, rng.normal(0, 0.3, nq), rng.normal(0, 0.3, nq)


def _run(fn, gpu, dem, H, Qxo, Qyo, nx, ny):
    import torch
    nq = (nx + 1) * (ny + 1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)  # noqa: E731
    Qx = torch.full((nq,), 7.0, dtype=torch.float64, device=gpu)
    Qy = torch.full((nq,), 7.0, dtype=torch.float64, device=gpu)
    dt = fn(dev(H), dev(dem), Qx, Qy, dev(Qxo), dev(Qyo), nx, ny, 5.0, 5.0, 0.035, 0.001, 10.0,
            0.01, 10.0, 1.0, 25.0, 0.7, 9.81, 10.0, 1.0, -9999.0)
    return dt, Qx.cpu().numpy(), Qy.cpu().numpy()


def test_noalloc_floodplainq_equals_workspace_variant(gpu):
    cases = [(11, 300, 97, 3.0), (12, 41, 23, 0.5)]   # second: smaller grid, lower H max
    dts = []
    for seed, nx, ny, hmax in cases:
        dem, H, Qxo, Qyo = _case(seed, nx, ny, hmax)
        dt_s, qx_s, qy_s = _run(K.FloodplainQ_GPU_noalloc, gpu, dem, H, Qxo, Qyo, nx, ny)
        dt_w, qx_w, qy_w = _run(K.FloodplainQ_GPU, gpu, dem, H, Qxo, Qyo, nx, ny)
        print(f"noalloc {nx}x{ny}: Tstep {dt_s!r} (workspace {dt_w!r})", flush=True)
        assert dt_s == dt_w
        assert np.array_equal(qx_s, qx_w) and np.array_equal(qy_s, qy_w)
        dts.append(dt_s)
    assert dts[1] > dts[0]   # the lower second maximum gave the longer step
