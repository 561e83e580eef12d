"""GPU paths against the committed golden vectors of the 256 x 256 plane
workload (FP64 oracle trajectory, tests/golden/plane256.npz)."""
import numpy as np
import pytest

from adaptflood_amd import generators, kernels
from adaptflood_amd.domain import Domain, make_tuning

from golden_data import load

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9   # m, max |dH|
FP32_TOL = 1e-3   # m, max |dH|
FP32_MASS = 1e-5  # relative volume error


def test_fused_fp64_snapshots_and_end(gpu):
    g = load()
    scn = generators.plane_case()
    d = Domain(scn, precision="fp64", tuning=make_tuning(dt_log_cap=1 << 12))
    done = 0
    for k in (1, 10, 100):
        d.advance(k - done, graph=False)
        done = k
        err = float(np.max(np.abs(d.read("H") - g[f"H_{k}"])))
        print(f"fp64 step {k}: max|dH|={err:.3e}", flush=True)
        assert err <= FP64_TOL
        np.testing.assert_array_equal(d.read("H"), g[f"H_{k}"])
        assert d.info()["t"] == pytest.approx(float(g[f"t_{k}"]), abs=1e-9)
    assert float(np.max(np.abs(d.read("Qx") - g["Qx_100"]))) <= 1e-8
    assert float(np.max(np.abs(d.read("Qy") - g["Qy_100"]))) <= 1e-8
    d.run(check_every=216, graph=True)
    info = d.info()
    assert info["steps"] == int(g["steps_end"])
    np.testing.assert_allclose(d.dt_log(), g["dt"], rtol=0, atol=1e-12)
    err = float(np.max(np.abs(d.read("H") - g["H_end"])))
    print(f"fp64 end: max|dH|={err:.3e}", flush=True)
    assert err <= FP64_TOL
    np.testing.assert_array_equal(d.read("H"), g["H_end"])
    np.testing.assert_array_equal(d.read("Qx"), g["Qx_end"])
    rep = d.reports()
    assert rep.shape == g["reports"].shape
    np.testing.assert_allclose(rep[:, 2], g["reports"][:, 2], rtol=1e-12)


def test_fused_fp32_end(gpu):
    g = load()
    scn = generators.plane_case()
    d = Domain(scn, precision="fp32")
    d.run(check_every=216, graph=True)
    H = d.read("H")
    err = float(np.max(np.abs(H - g["H_end"])))
    rms = float(np.sqrt(np.mean((H - g["H_end"]) ** 2)))
    vol, _ = d.volume()
    vref = float(np.sum(g["H_end"])) * scn.dx * scn.dx
    rel = abs(vol - vref) / vref
    print(f"fp32 end: max|dH|={err:.3e} rms={rms:.3e} rel mass={rel:.3e}", flush=True)
    assert err <= FP32_TOL
    assert rel <= FP32_MASS


def test_compat_sequence_end(gpu):
    """The reference call sequence through the per-function C ABI."""
    g = load()
    r = kernels.run_sequence(generators.plane_case())
    assert r["steps"] == int(g["steps_end"])
    np.testing.assert_allclose(np.array(r["dt"]), g["dt"], rtol=0, atol=1e-12)
    err = float(np.max(np.abs(r["H"] - g["H_end"])))
    print(f"compat end: max|dH|={err:.3e}", flush=True)
    assert err <= FP64_TOL
    np.testing.assert_array_equal(r["H"], g["H_end"])
    np.testing.assert_array_equal(r["Qy"], g["Qy_end"])
    np.testing.assert_allclose(np.array(r["reports"])[:, 2], g["reports"][:, 2], rtol=1e-12)
