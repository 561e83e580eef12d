"""The hand-over of tests/golden/make_full_run_digests.py --from-fixture: an oracle
run resumed from a partial fixture and a state whose digests equal its last
checkpoint continues exactly as the run that never stopped (same state, dt
log, clock and mass reports); a state that differs in one bit is refused."""
import importlib.util
import os

import numpy as np
import pytest

from adaptflood_amd import generators
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _digests():
    spec = importlib.util.spec_from_file_location(
        "make_full_run_digests", os.path.join(HERE, "golden", "make_full_run_digests.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _partial_fixture(m, r):
    st, rec = m.record(r)
    return {"checkpoints": [rec], "final": None, "dt_steps": len(st["dt"]),
            "dt_sha256": m.sha(st["dt"]), "reports": st["reports"].tolist()}


def test_resume_from_digests_continues_the_run():
    m = _digests()
    scn = generators.valley_case(n=128, n_qvar=40)
    scn.mass_int = 2.0      # several reports before and after the hand-over
    a = oracle.OracleRun(scn, dt_log_cap=1 << 12, mass_log_cap=256)
    a.advance(60)
    fx = _partial_fixture(m, a)
    st = a.state()
    b = oracle.OracleRun(scn, dt_log_cap=1 << 12, mass_log_cap=256)
    cps = m.resume(b, fx, st["H"].copy(), st["Qx"].copy(), st["Qy"].copy(), st["dt"].copy())
    assert cps == fx["checkpoints"] and b.steps == 60 and b.t == a.t
    assert b.next_mass == a.next_mass and len(b.reports) >= 3
    a.advance(90)
    b.advance(90)
    sa, sb = a.state(), b.state()
    for k in ("H", "Qx", "Qy", "Qxold", "Qyold", "dt", "reports"):
        assert np.array_equal(sa[k], sb[k]), k
    assert sa["t"] == sb["t"] and sa["steps"] == sb["steps"] == 150


def test_resume_refuses_a_state_off_by_one_bit():
    m = _digests()
    scn = generators.valley_case(n=128, n_qvar=40)
    a = oracle.OracleRun(scn, dt_log_cap=1 << 12, mass_log_cap=256)
    a.advance(30)
    fx = _partial_fixture(m, a)
    st = a.state()
    H = st["H"].copy()
    k = int(np.argmax(H))
    H.ravel()[k] = np.nextafter(H.ravel()[k], np.inf)
    b = oracle.OracleRun(scn, dt_log_cap=1 << 12, mass_log_cap=256)
    with pytest.raises(ValueError, match="H at step 30"):
        m.resume(b, fx, H, st["Qx"], st["Qy"], st["dt"])
    dts = st["dt"].copy()
    dts[-1] = np.nextafter(dts[-1], 0)
    with pytest.raises(ValueError, match="dt log"):
        m.resume(b, fx, st["H"], st["Qx"], st["Qy"], dts)
