// adaptflood_compat.hpp -- C++ wrappers with the reference's function names
// and void signatures (src/kernels.hpp:8-37) over the C ABI of adaptflood.h.
//
// A reference caller swaps `#include "kernels.hpp"` for this header and
// passes device pointers (`view.data()`) where it passed Kokkos views.  Each
// wrapper returns only when its work is done, as the reference's functions
// do (they end in Kokkos::fence), and throws af::compat::error where the
// reference would have aborted.
//
// Per calling thread: the HIP stream the calls go to (default: the null
// stream) and the device workspace FloodplainQ_GPU reduces H-max into
// (AF_FLOODPLAINQ_WORKSPACE_BYTES, owned by the caller; the reference let
// Kokkos allocate it implicitly).  Set both with af::compat::bind().
#pragma once
#include <stdexcept>
#include <string>

#include "adaptflood.h"

namespace af {
namespace compat {

struct error : std::runtime_error {
    int code;
    error(const char* what, int rc)
        : std::runtime_error(std::string(what) + ": " + af_last_error()), code(rc) {}
};

inline void check(int rc, const char* what) {
    if (rc != AF_OK) throw error(what, rc);
}

struct binding {
    af_stream_t stream = nullptr;
    void* workspace = nullptr;
};

inline binding& current() {
    static thread_local binding b;
    return b;
}

// Binds the calling thread's stream and FloodplainQ workspace.
inline void bind(af_stream_t stream, void* workspace) {
    current().stream = stream;
    current().workspace = workspace;
}

}  // namespace compat
}  // namespace af

// src/kernels.hpp:8-9
inline void UpdateQs_GPU(const double* Qx, const double* Qy, double* Qxold, double* Qyold, int xsz,
                         int ysz, double dx) {
    af::compat::check(af_UpdateQs_GPU(Qx, Qy, Qxold, Qyold, xsz, ysz, dx,
                                      af::compat::current().stream),
                      "UpdateQs_GPU");
}

// src/kernels.hpp:12-16; Tstep is the out-parameter, as in the reference
inline void FloodplainQ_GPU(const double* H, const double* DEM, double* Qx, double* Qy,
                            const double* Qxold, const double* Qyold, int xsz, int ysz, double dx,
                            double dy, double FPn, double DepthThresh, double MaxHflow,
                            double dhlin, double& Tstep, double Qlimfact, double dA, double cfl,
                            double g, double InitTstep, double theta, double nodata) {
    af::compat::check(
        af_FloodplainQ_GPU(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, dy, FPn, DepthThresh,
                           MaxHflow, dhlin, &Tstep, Qlimfact, dA, cfl, g, InitTstep, theta, nodata,
                           af::compat::current().workspace, af::compat::current().stream),
        "FloodplainQ_GPU");
}

// src/kernels.hpp:19-21
inline void BCs_GPU(double* Qx, double* Qy, const double* H, const double* DEM, int xsz, int ysz,
                    double dx, double g, double dt, double nodata, bool west_free, bool east_free,
                    bool north_free, bool south_free) {
    af::compat::check(af_BCs_GPU(Qx, Qy, H, DEM, xsz, ysz, dx, g, dt, nodata, west_free,
                                 east_free, north_free, south_free, af::compat::current().stream),
                      "BCs_GPU");
}

// src/kernels.hpp:24-25
inline void ApplyPointBCs_GPU(double* H, const int32_t* bc_idx, const int32_t* bc_type,
                              const double* bc_value, int num_bcs, int ncols, double dx, double dA,
                              double Tstep) {
    af::compat::check(af_ApplyPointBCs_GPU(H, bc_idx, bc_type, bc_value, num_bcs, ncols, dx, dA,
                                           Tstep, af::compat::current().stream),
                      "ApplyPointBCs_GPU");
}

// src/kernels.hpp:28
inline void UpdateH_GPU(double* H, const double* Qx, const double* Qy, int xsz, int ysz, double dA,
                        double Tstep) {
    af::compat::check(
        af_UpdateH_GPU(H, Qx, Qy, xsz, ysz, dA, Tstep, af::compat::current().stream),
        "UpdateH_GPU");
}

// src/kernels.hpp:31-33
inline void ComputeVelocity_GPU(double* Vx_out, double* Vy_out, const double* H, const double* DEM,
                                const double* Qx, const double* Qy, int xsz, int ysz, double dx,
                                double nodata) {
    af::compat::check(af_ComputeVelocity_GPU(Vx_out, Vy_out, H, DEM, Qx, Qy, xsz, ysz, dx, nodata,
                                             af::compat::current().stream),
                      "ComputeVelocity_GPU");
}

// src/kernels.hpp:36-37
inline void Rainfall_GPU(double* H, const double* DEM, int xsz, int ysz, double rain_rate,
                         double Tstep, double nodata) {
    af::compat::check(af_Rainfall_GPU(H, DEM, xsz, ysz, rain_rate, Tstep, nodata,
                                      af::compat::current().stream),
                      "Rainfall_GPU");
}

// The source injection the reference writes inline (src/main.cpp:256-260).
inline void AddFlowBCs_GPU(double* H, const int32_t* ps_idx, const double* ps_flow,
                           int num_flow_bcs) {
    af::compat::check(
        af_AddFlowBCs_GPU(H, ps_idx, ps_flow, num_flow_bcs, af::compat::current().stream),
        "AddFlowBCs_GPU");
}
