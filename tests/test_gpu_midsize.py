"""Parity on the generated workloads at 1024 x 1024: the valley case (HFIX
west edge, 2000 QVAR hydrographs, rain, fractal terrain with a carved valley)
and the roofline case (NODATA blobs, 5000 QFIX, free edges), 100 steps each.
The FP64 fused path must equal the oracle bitwise, state, dt sequence and
fluxes; the FP32 path stays within its bar over the first 20 steps (these
flows amplify differences; later steps are measured by the accuracy
workload, not asserted)."""
import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd.domain import Domain, make_tuning
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {"valley": generators.valley_case, "roofline": generators.roofline_case}


@pytest.mark.parametrize("case", ["valley", "roofline"])
def test_fp64_midsize_bitwise(gpu, case):
    scn = CASES[case](n=1024)
    ref = oracle.run(scn, max_steps=100)
    d = Domain(scn, precision="fp64", tuning=make_tuning(dt_log_cap=256))
    d.advance(100, graph=True)
    H = d.read("H")
    print(f"{case} fp64: t={d.info()['t']:.3f} (oracle {ref['t']:.3f}) "
          f"max|dH|={float(np.max(np.abs(H - ref['H']))):.2e}", flush=True)
    assert d.info()["steps"] == 100 and d.info()["t"] == ref["t"]
    np.testing.assert_array_equal(d.dt_log(), ref["dt"])
    np.testing.assert_array_equal(H, ref["H"])
    np.testing.assert_array_equal(d.read("Qx"), ref["Qx"])
    np.testing.assert_array_equal(d.read("Qy"), ref["Qy"])


@pytest.mark.parametrize("case", ["valley", "roofline"])
def test_fp32_midsize_early_steps(gpu, case):
    scn = CASES[case](n=1024)
    ref = oracle.run(scn, max_steps=20)
    d = Domain(scn, precision="fp32")
    d.advance(20, graph=False)
    H = d.read("H")
    err = float(np.max(np.abs(H - ref["H"])))
    vol, _ = d.volume()
    vref = float(np.sum(ref["H"])) * scn.dx * scn.dx
    print(f"{case} fp32 20 steps: max|dH|={err:.2e} rel vol={abs(vol - vref) / vref:.2e}",
          flush=True)
    assert err <= 1e-3
    assert abs(vol - vref) / vref <= 1e-5
