"""The reference call surface (af_*_GPU, FP64, reference layout) vs the oracle (GPU).

Every function is compared one at a time on random inputs, then the whole
main.cpp sequence composed from them is compared over full runs.
"""
import numpy as np
import pytest

from adaptflood_amd import generators, kernels as K
from oracle import oracle

pytestmark = pytest.mark.gpu


def _dev(a, gpu, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(gpu)


def _state(seed, nx, ny):
    rng = np.random.default_rng(seed)
    dem = rng.uniform(0, 3, nx * ny)
    dem[rng.random(nx * ny) < 0.08] = -9999.0
    H = np.maximum(rng.uniform(-0.2, 1.0, nx * ny), 0.0)
    nq = (nx + 1) * (ny + 1)
    Qxo = rng.normal(0, 0.3, nq)
    Qyo = rng.normal(0, 0.3, nq)
    return dem, H, Qxo, Qyo


@pytest.mark.parametrize("seed,nx,ny", [(1, 31, 17), (2, 64, 64), (3, 200, 9)])
def test_floodplainq_and_bcs(gpu, seed, nx, ny):
    import torch
    dem, H, Qxo, Qyo = _state(seed, nx, ny)
    nq = (nx + 1) * (ny + 1)
    for theta in (1.0, 0.7):
        Qx = np.zeros(nq)
        Qy = np.zeros(nq)
        dt_ref = oracle.floodplain_q(H, dem, Qx, Qy, Qxo, Qyo, nx, ny, 5.0, 0.035, 0.001, 10.0,
                                     0.7, 9.81, 10.0, theta, -9999.0)
        gQx = torch.zeros(nq, dtype=torch.float64, device=gpu)
        gQy = torch.zeros(nq, dtype=torch.float64, device=gpu)
        gH, gD = _dev(H, gpu), _dev(dem, gpu)
        dt = K.FloodplainQ_GPU(gH, gD, gQx, gQy, _dev(Qxo, gpu), _dev(Qyo, gpu), nx, ny, 5.0, 5.0,
                               0.035, 0.001, 10.0, 0.01, 10.0, 1.0, 25.0, 0.7, 9.81, 10.0, theta,
                               -9999.0)
        assert dt == dt_ref
        np.testing.assert_allclose(gQx.cpu().numpy(), Qx, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(gQy.cpu().numpy(), Qy, rtol=1e-13, atol=1e-15)
        P = nx + 1
        for flags in ((1, 1, 1, 1), (0, 0, 0, 0), (1, 0, 0, 1)):
            oracle.bcs(Qx, Qy, H, dem, nx, ny, 5.0, 9.81, -9999.0, *flags)
            K.BCs_GPU(gQx, gQy, gH, gD, nx, ny, 5.0, 9.81, dt, -9999.0, *flags)
            gx = gQx.cpu().numpy().reshape(ny + 1, P)
            gy = gQy.cpu().numpy().reshape(ny + 1, P)
            rx, ry = Qx.reshape(ny + 1, P), Qy.reshape(ny + 1, P)
            # edge faces are sqrt/multiply only: bitwise
            assert np.array_equal(gx[:, 0], rx[:, 0]) and np.array_equal(gx[:, nx], rx[:, nx])
            assert np.array_equal(gy[0, :], ry[0, :]) and np.array_equal(gy[ny, :], ry[ny, :])
            np.testing.assert_allclose(gx, rx, rtol=1e-13, atol=1e-15)
            np.testing.assert_allclose(gy, ry, rtol=1e-13, atol=1e-15)


def test_update_rain_points_qs(gpu):
    import torch
    nx, ny = 53, 41
    dem, H, Qx, Qy = _state(9, nx, ny)
    dA, dt = 25.0, 1.3
    gH = _dev(H, gpu)
    gD = _dev(dem, gpu)
    # AddFlowBCs with a duplicated cell, then rain, update, overrides, q_old
    idx = np.array([5, 77, 5, 900, 77], dtype=np.int32)
    ps = np.array([0.2, -0.5, 0.1, 1e-3, 0.3])
    oracle.add_flow_bcs(H, idx, ps)
    K.AddFlowBCs_GPU(gH, _dev(idx, gpu), _dev(ps, gpu), len(idx))
    assert np.array_equal(gH.cpu().numpy(), H)
    oracle.rainfall(H, dem, nx, ny, 2e-5, dt, -9999.0)
    K.Rainfall_GPU(gH, gD, nx, ny, 2e-5, dt, -9999.0)
    assert np.array_equal(gH.cpu().numpy(), H)
    oracle.update_h(H, Qx, Qy, nx, ny, dA, dt)
    K.UpdateH_GPU(gH, _dev(Qx, gpu), _dev(Qy, gpu), nx, ny, dA, dt)
    assert np.array_equal(gH.cpu().numpy(), H)
    bidx = np.array([3, 3, 60, 61], dtype=np.int32)
    code = np.array([2, 1, 3, 2], dtype=np.int32)
    val = np.array([0.4, 0.0, 0.25, -1.0])
    oracle.apply_point_bcs(H, bidx, code, val)
    K.ApplyPointBCs_GPU(gH, _dev(bidx, gpu), _dev(code, gpu), _dev(val, gpu), len(bidx), nx, 5.0,
                        dA, dt)
    assert np.array_equal(gH.cpu().numpy(), H)
    nq = (nx + 1) * (ny + 1)
    Qxo, Qyo = np.zeros(nq), np.zeros(nq)
    oracle.update_qs(Qx, Qy, Qxo, Qyo, nx, ny, 5.0)
    gxo = torch.zeros(nq, dtype=torch.float64, device=gpu)
    gyo = torch.zeros(nq, dtype=torch.float64, device=gpu)
    K.UpdateQs_GPU(_dev(Qx, gpu), _dev(Qy, gpu), gxo, gyo, nx, ny, 5.0)
    assert np.array_equal(gxo.cpu().numpy(), Qxo)
    assert np.array_equal(gyo.cpu().numpy(), Qyo)
    Vx, Vy = oracle.velocity(H, dem, Qx, Qy, nx, ny, 5.0, -9999.0)
    gvx = torch.zeros(nx * ny, dtype=torch.float64, device=gpu)
    gvy = torch.zeros(nx * ny, dtype=torch.float64, device=gpu)
    K.ComputeVelocity_GPU(gvx, gvy, gH, gD, _dev(Qx, gpu), _dev(Qy, gpu), nx, ny, 5.0, -9999.0)
    assert np.array_equal(gvx.cpu().numpy(), Vx)
    assert np.array_equal(gvy.cpu().numpy(), Vy)


def test_sequence_small(gpu):
    scn = generators.small_case(48, 36, seed=5, sim_time=150.0)
    ref = oracle.run(scn)
    out = K.run_sequence(scn)
    assert out["steps"] == ref["steps"]
    np.testing.assert_allclose(out["dt"], ref["dt"], rtol=0, atol=1e-12)
    assert float(np.max(np.abs(out["H"] - ref["H"]))) <= 1e-9


def test_sequence_plane_case(gpu):
    scn = generators.plane_case(sim_time=600.0)
    ref = oracle.run(scn)
    out = K.run_sequence(scn)
    assert out["steps"] == ref["steps"]
    err = float(np.max(np.abs(out["H"] - ref["H"])))
    print(f"compat sequence plane 600 s: steps={out['steps']} max|dH|={err:.3e}", flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("n", [1, 2, 5000, 8192, 9000, 20000])
def test_point_lists_with_duplicates(gpu, n):
    """AddFlowBCs_GPU and ApplyPointBCs_GPU on long lists (chunks of up to
    8192 entries, each grouped by one sort in LDS, taken in array order; a
    cell's entries fall in several chunks when the list is longer), with
    cells repeated several times and negative depths to clamp: bitwise the
    oracle's sequential file-order loops (main.cpp:256-260, kernels.cpp:261-274)."""
    nx, ny = 300, 211
    rng = np.random.default_rng(1000 + n)
    dem, H, _, _ = _state(1000 + n, nx, ny)
    cells = rng.integers(0, nx * ny, max(1, n // 3))
    idx = rng.choice(cells, n).astype(np.int32)          # many duplicates
    ps = rng.normal(0.0, 0.3, n)
    Href = H.copy()
    oracle.add_flow_bcs(Href, idx, ps)
    gH = _dev(H, gpu)
    K.AddFlowBCs_GPU(gH, _dev(idx, gpu), _dev(ps, gpu), n)
    assert np.array_equal(gH.cpu().numpy(), Href)
    code = rng.integers(1, 4, n).astype(np.int32)
    code[rng.random(n) < 0.1] = 4                         # not an override: ignored
    val = rng.uniform(0.0, 2.0, n)
    oracle.apply_point_bcs(Href, idx, code, val)
    K.ApplyPointBCs_GPU(gH, _dev(idx, gpu), _dev(code, gpu), _dev(val, gpu), n, nx, 5.0, 25.0,
                        1.0)
    assert np.array_equal(gH.cpu().numpy(), Href)
