"""The FP64 fused path against the FP64 oracle's full runs, through SHA-256
digests: the 4096 x 4096 valley workload (fractal terrain with a carved
valley at 5 m, 2000 QVAR hydrographs, rain over 0-1800 s, HFIX 2 m west
edge, FREE N/E/S; 7200 s) and the roofline generator at 1024 x 1024
(fractal terrain at 2 m, 10% NODATA, 5000 QFIX, FREE edges; 3600 s).  The
fixtures (tests/golden/<case>_fp64.json, written by
tests/golden/make_valley_digests.py from the oracle) keep digests of H, Qx
and Qy every 5000 steps and at the end of the run, the digest of the dt log,
and the mass reports.  The fused path must match every digest, so its
state, fluxes and step sizes equal the oracle's bitwise over the whole run.
Report times, steps and wet areas are exact; the volumes agree to n*eps
(n = cells): the oracle sums them with a plain threaded reduction whose
rounding error can reach (n-1)*eps relative, the fused path with
compensated (Kahan) partial sums."""
import hashlib
import json
import os

import numpy as np
import pytest

from adaptflood_amd import generators
from adaptflood_amd.domain import Domain, make_tuning

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"valley4096": lambda: generators.valley_case(n=4096),
         "roofline1024": lambda: generators.roofline_case(n=1024)}


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_matches_oracle_digests(gpu, case):
    with open(os.path.join(GOLDEN, f"{case}_fp64.json")) as f:
        fx = json.load(f)
    scn = CASES[case]()
    assert hashlib.sha256(np.ascontiguousarray(scn.dem, dtype=np.float64).tobytes()
                          ).hexdigest() == fx["dem_sha256"]
    records = list(fx["checkpoints"]) + ([fx["final"]] if fx["final"] else [])
    d = Domain(scn, precision="fp64",
               tuning=make_tuning(dt_log_cap=fx["dt_steps"] + 64, report_cap=1024))
    done = 0
    for rec in records:
        d.advance(rec["steps"] - done, graph=True)
        done = rec["steps"]
        info = d.info()
        H = d.read("H")
        print(f"{case} fp64 step {info['steps']} t={info['t']:.6f} (oracle {rec['t']:.6f}) "
              f"H max {float(H.max()):.6f}", flush=True)
        assert info["steps"] == rec["steps"] and info["t"] == rec["t"]
        assert _sha(H) == rec["H"], rec["steps"]
        assert _sha(d.read("Qx")) == rec["Qx"], rec["steps"]
        assert _sha(d.read("Qy")) == rec["Qy"], rec["steps"]
    if fx["final"]:
        assert d.info()["finished"]
    dts = d.dt_log()
    assert _sha(dts[: fx["dt_steps"]]) == fx["dt_sha256"]
    ref = np.array(fx["reports"]).reshape(-1, 4)
    got = d.reports()[: len(ref)]
    assert len(got) == len(ref)
    assert np.array_equal(got[:, [0, 1, 3]], ref[:, [0, 1, 3]])
    n_eps = scn.cells * np.finfo(np.float64).eps / 2
    np.testing.assert_allclose(got[:, 2], ref[:, 2], rtol=n_eps, atol=0)
