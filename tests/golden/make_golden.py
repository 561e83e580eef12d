"""Writes the golden fixtures of the 256 x 256 plane workload (a QFIX source
of 5 m^3/s at the centre cell, FREE east edge, 3600 s) from the FP64 oracle.

The oracle is pinned by tests/test_oracle.py (closed forms, known answers and
invariants); these vectors freeze its trajectory so the GPU paths and later
oracle changes are checked against the same numbers without re-running it.
Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from adaptflood_amd import generators  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plane256.npz")


def dem_digest(dem):
    return hashlib.sha256(np.ascontiguousarray(dem, dtype=np.float64).tobytes()).hexdigest()


def main():
    scn = generators.plane_case()
    r = oracle.OracleRun(scn)
    snaps = {}
    done = 0
    for k in (1, 10, 100):
        r.advance(k - done)
        done = k
        st = r.state()
        snaps[f"H_{k}"] = st["H"].copy()
        snaps[f"t_{k}"] = np.array(st["t"])
        if k == 100:
            snaps["Qx_100"] = st["Qx"].copy()
            snaps["Qy_100"] = st["Qy"].copy()
    r.advance()
    st = r.state()
    np.savez_compressed(
        OUT, dem_sha256=np.array(dem_digest(scn.dem)), H_end=st["H"], Qx_end=st["Qx"],
        Qy_end=st["Qy"], t_end=np.array(st["t"]), steps_end=np.array(st["steps"]),
        dt=st["dt"], reports=st["reports"], **snaps)
    print(f"wrote {OUT}: {st['steps']} steps, t={st['t']:.3f}, "
          f"{os.path.getsize(OUT) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
