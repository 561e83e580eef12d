"""The reference's step-kernel call surface (src/kernels.hpp:8-37) over the C ABI.

Each function keeps the reference's name, argument order and meaning.  Views
become torch tensors on the GPU (float64 for RealView, int32 for IntView);
`Tstep` of FloodplainQ_GPU, an out-parameter in C++, is returned.  Every call
runs HIP kernels of libadaptflood_amd.so and returns after they finish, as
the reference's functions do (each ends in Kokkos::fence()).
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import call


def _p(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device tensor expected")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    return C.c_void_p(t.data_ptr())


def _s(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _f64(*ts):
    import torch
    for t in ts:
        if t.dtype != torch.float64:
            raise TypeError("RealView arguments must be float64 tensors")


def _i32(*ts):
    import torch
    for t in ts:
        if t.dtype != torch.int32:
            raise TypeError("IntView arguments must be int32 tensors")


_workspaces = {}


def _workspace(device):
    import torch
    key = str(device)
    ws = _workspaces.get(key)
    if ws is None:
        ws = torch.empty(_lib.AF_FLOODPLAINQ_WORKSPACE_BYTES, dtype=torch.uint8, device=device)
        _workspaces[key] = ws
    return ws


def UpdateQs_GPU(Qx, Qy, Qxold, Qyold, xsz, ysz, dx, stream=None):
    """kernels.cpp:6-19"""
    _f64(Qx, Qy, Qxold, Qyold)
    call("af_UpdateQs_GPU", _p(Qx), _p(Qy), _p(Qxold), _p(Qyold), int(xsz), int(ysz), float(dx),
         _s(stream))


def FloodplainQ_GPU(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, dy, FPn, DepthThresh,
                    MaxHflow, dhlin, Tstep, Qlimfact, dA, cfl, g, InitTstep, theta, nodata,
                    stream=None):
    """kernels.cpp:22-146.  Returns the new Tstep (the reference's out-parameter;
    the value passed in is ignored, as the reference overwrites it)."""
    _f64(H, DEM, Qx, Qy, Qxold, Qyold)
    ts = C.c_double(float(Tstep))
    ws = _workspace(H.device)
    call("af_FloodplainQ_GPU", _p(H), _p(DEM), _p(Qx), _p(Qy), _p(Qxold), _p(Qyold), int(xsz),
         int(ysz), float(dx), float(dy), float(FPn), float(DepthThresh), float(MaxHflow),
         float(dhlin), C.byref(ts), float(Qlimfact), float(dA), float(cfl), float(g),
         float(InitTstep), float(theta), float(nodata), _p(ws), _s(stream))
    return ts.value


def FloodplainQ_GPU_noalloc(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, dy, FPn, DepthThresh,
                            MaxHflow, dhlin, Tstep, Qlimfact, dA, cfl, g, InitTstep, theta,
                            nodata, stream=None):
    """FloodplainQ_GPU with its H-max key in the library's static device
    scratch instead of a workspace (no counterpart in the reference; for
    callers that cannot allocate).  Calls on one GPU must not overlap."""
    _f64(H, DEM, Qx, Qy, Qxold, Qyold)
    ts = C.c_double(float(Tstep))
    call("af_FloodplainQ_GPU_noalloc", _p(H), _p(DEM), _p(Qx), _p(Qy), _p(Qxold), _p(Qyold),
         int(xsz), int(ysz), float(dx), float(dy), float(FPn), float(DepthThresh),
         float(MaxHflow), float(dhlin), C.byref(ts), float(Qlimfact), float(dA), float(cfl),
         float(g), float(InitTstep), float(theta), float(nodata), _s(stream))
    return ts.value


def BCs_GPU(Qx, Qy, H, DEM, xsz, ysz, dx, g, dt, nodata, west_free, east_free, north_free,
            south_free, stream=None):
    """kernels.cpp:149-256"""
    _f64(Qx, Qy, H, DEM)
    call("af_BCs_GPU", _p(Qx), _p(Qy), _p(H), _p(DEM), int(xsz), int(ysz), float(dx), float(g),
         float(dt), float(nodata), int(bool(west_free)), int(bool(east_free)),
         int(bool(north_free)), int(bool(south_free)), _s(stream))


def ApplyPointBCs_GPU(H, bc_idx, bc_type, bc_value, num_bcs, ncols, dx, dA, Tstep, stream=None):
    """kernels.cpp:259-277"""
    _f64(H, bc_value)
    _i32(bc_idx, bc_type)
    call("af_ApplyPointBCs_GPU", _p(H), _p(bc_idx), _p(bc_type), _p(bc_value), int(num_bcs),
         int(ncols), float(dx), float(dA), float(Tstep), _s(stream))


def UpdateH_GPU(H, Qx, Qy, xsz, ysz, dA, Tstep, stream=None):
    """kernels.cpp:280-290"""
    _f64(H, Qx, Qy)
    call("af_UpdateH_GPU", _p(H), _p(Qx), _p(Qy), int(xsz), int(ysz), float(dA), float(Tstep),
         _s(stream))


def Rainfall_GPU(H, DEM, xsz, ysz, rain_rate, Tstep, nodata, stream=None):
    """kernels.cpp:354-367"""
    _f64(H, DEM)
    call("af_Rainfall_GPU", _p(H), _p(DEM), int(xsz), int(ysz), float(rain_rate), float(Tstep),
         float(nodata), _s(stream))


def AddFlowBCs_GPU(H, ps_idx, ps_flow, num_flow_bcs, stream=None):
    """main.cpp:256-260 (inline in the reference)"""
    _f64(H, ps_flow)
    _i32(ps_idx)
    call("af_AddFlowBCs_GPU", _p(H), _p(ps_idx), _p(ps_flow), int(num_flow_bcs), _s(stream))


def ComputeVelocity_GPU(Vx_out, Vy_out, H, DEM, Qx, Qy, xsz, ysz, dx, nodata, stream=None):
    """kernels.cpp:293-351"""
    _f64(Vx_out, Vy_out, H, DEM, Qx, Qy)
    call("af_ComputeVelocity_GPU", _p(Vx_out), _p(Vy_out), _p(H), _p(DEM), _p(Qx), _p(Qy),
         int(xsz), int(ysz), float(dx), float(nodata), _s(stream))


def interpolate_bc_value(times, values, value, t):
    """io.cpp:299-311 (host)"""
    n = 0 if times is None else len(times)
    if n == 0:
        return float(value)
    ts = np.ascontiguousarray(times, dtype=np.float64)
    vs = np.ascontiguousarray(values, dtype=np.float64)
    return _lib.lib().af_interp_series(_lib.dptr(ts), _lib.dptr(vs), n, float(value), float(t))


def run_sequence(scn, max_steps=None, device="cuda"):
    """The time loop of main.cpp:196-304 driven through the per-function C ABI
    above (FP64, reference layout): the drop-in composition of the reference's
    own call sequence.  Returns the final state, t, the step count, the dt of
    every step and the mass reports."""
    import torch
    nx, ny, dx = scn.nx, scn.ny, float(scn.dx)
    dA = dx * dx
    dev = torch.device(device)
    H = torch.zeros(nx * ny, dtype=torch.float64, device=dev)
    if scn.h0 is not None:
        H.copy_(torch.from_numpy(np.ascontiguousarray(scn.h0, dtype=np.float64).ravel()))
    DEM = torch.from_numpy(np.ascontiguousarray(scn.dem, dtype=np.float64).ravel()).to(dev)
    nq = (nx + 1) * (ny + 1)
    Qx, Qy, Qxold, Qyold = (torch.zeros(nq, dtype=torch.float64, device=dev) for _ in range(4))
    flow = [p for p in scn.points if p.type in (_lib.AF_BC_QFIX, _lib.AF_BC_QVAR)]
    other = [p for p in scn.points if p.type in (_lib.AF_BC_FREE, _lib.AF_BC_HFIX,
                                                _lib.AF_BC_HVAR)]
    ps_idx = torch.tensor([p.x + p.y * nx for p in flow] or [0], dtype=torch.int32, device=dev)
    bc_idx = torch.tensor([p.x + p.y * nx for p in other] or [0], dtype=torch.int32, device=dev)
    code = {_lib.AF_BC_FREE: 1, _lib.AF_BC_HFIX: 2, _lib.AF_BC_HVAR: 3}
    bc_type = torch.tensor([code[p.type] for p in other] or [0], dtype=torch.int32, device=dev)
    bc_value_host = np.array([p.value for p in other] or [0.0], dtype=np.float64)
    rain = scn.rain_times is not None and len(scn.rain_times) > 0
    t, step = 0.0, 0
    torch.cuda.synchronize()
    t_loop = time.perf_counter()
    next_mass = scn.mass_int
    dts, reports = [], []
    while t < scn.sim_time:
        if max_steps is not None and step >= max_steps:
            break
        Tstep = FloodplainQ_GPU(H, DEM, Qx, Qy, Qxold, Qyold, nx, ny, dx, dx, scn.fpn,
                                scn.depth_thresh, scn.max_hflow, 0.01, scn.init_tstep, 1.0, dA,
                                scn.cfl, scn.g, scn.init_tstep, scn.theta, scn.nodata)
        BCs_GPU(Qx, Qy, H, DEM, nx, ny, dx, scn.g, Tstep, scn.nodata, scn.west_free,
                scn.east_free, scn.north_free, scn.south_free)
        if flow:
            ps = np.empty(len(flow))
            for s, p in enumerate(flow):
                v = p.value if p.type == _lib.AF_BC_QFIX else interpolate_bc_value(
                    p.times, p.values, p.value, t)
                ps[s] = v * dx * Tstep / dA
            AddFlowBCs_GPU(H, ps_idx, torch.from_numpy(ps).to(dev), len(flow))
        if rain:
            r = interpolate_bc_value(scn.rain_times, scn.rain_rates, 0.0, t)
            Rainfall_GPU(H, DEM, nx, ny, r, Tstep, scn.nodata)
        UpdateH_GPU(H, Qx, Qy, nx, ny, dA, Tstep)
        if other:
            for b, p in enumerate(other):
                if p.type == _lib.AF_BC_HVAR:
                    bc_value_host[b] = interpolate_bc_value(p.times, p.values, p.value, t)
            ApplyPointBCs_GPU(H, bc_idx, bc_type, torch.from_numpy(bc_value_host).to(dev),
                              len(other), nx, dx, dA, Tstep)
        UpdateQs_GPU(Qx, Qy, Qxold, Qyold, nx, ny, dx)
        dts.append(Tstep)
        t += Tstep
        step += 1
        if t >= next_mass:  # main.cpp:291-304, reduced on the device
            vol = float((H * dA).sum().item())
            area = float((H > 0.01).sum().item()) * dA
            reports.append((t, Tstep, vol, area))
            next_mass += scn.mass_int
    torch.cuda.synchronize()
    loop_seconds = time.perf_counter() - t_loop
    return {"loop_seconds": loop_seconds,
            "H": H.cpu().numpy().reshape(ny, nx), "Qx": Qx.cpu().numpy(), "Qy": Qy.cpu().numpy(),
            "Qxold": Qxold.cpu().numpy(), "Qyold": Qyold.cpu().numpy(), "t": t, "steps": step,
            "dt": np.array(dts), "reports": np.array(reports).reshape(-1, 4)}
