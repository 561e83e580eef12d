// af_device.hpp -- device arithmetic shared by the per-function kernels and
// the fused step kernel.  All of it is compiled with -ffp-contract=off so the
// FP64 forms evaluate exactly the operations the reference evaluates, in the
// reference's association order (src/kernels.cpp:58-92, 108-142, 149-256,
// 280-290, 354-367).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace af {

// types.hpp:14-15
template <typename T>
__host__ __device__ __forceinline__ T getmax(T a, T b) { return (a > b) ? a : b; }
template <typename T>
__host__ __device__ __forceinline__ T getmin(T a, T b) { return (a < b) ? a : b; }

// pow(h, 10.0/3.0) for h > 0 (the friction term of kernels.cpp:83 and :133),
// built from IEEE operations only (+, -, *, fma, frexp, ldexp), so it
// rounds the same way on every machine that rounds those correctly.
//   h = xr * 2^(3q), xr in [1, 8):  h^(10/3) = xr^3 * cbrt(xr) * 2^(10q).
//   cbrt(xr): Newton's iteration for the inverse cube root y = xr^(-1/3),
//     y <- y (4 - xr y^3) / 3, from a quadratic start (6 steps), then one
//     Newton correction of c0 = xr y^2 with the residual xr - c0^3 taken in
//     double-double (Dekker's exact products through fma).
//   xr^3 in double-double, times cbrt as a double-double product.
//   The exponent 10.0/3.0 exceeds 10/3 by 1.4802973661668753e-16; that
//     factor h^1.48e-16 = 1 + 1.48e-16 ln h is applied to first order (ln h
//     from a cubic fit on [1, 2) is ample: the term is below 5 ulp).
// The result is the correctly rounded pow(h, 10.0/3.0) but for about 3 in
// 10^4 arguments, which are 1 ulp away; it differs from glibc's pow in about
// 1 in 10^3 arguments, by 1 ulp.
__host__ __device__ __forceinline__ double pow_10_3(double h) {
    int E;
    const double f = frexp(h, &E);                       // h = f 2^E, f in [0.5, 1)
    const int e = E - 1;
    const int q = (e >= 0) ? e / 3 : -((2 - e) / 3);     // floor(e / 3)
    const double xr = ldexp(2.0 * f, e - 3 * q);         // [1, 8)
    double y = 1.190645532750796 + xr * (-0.20554765291607405 + xr * 0.01490212016527807);
    for (int i = 0; i < 6; i++) y = y * (4.0 - xr * (y * y * y)) * (1.0 / 3.0);
    const double c0 = xr * (y * y);
    const double p = c0 * c0, pe = fma(c0, c0, -p);
    const double t = p * c0, te = fma(p, c0, -t) + pe * c0;
    const double c1 = ((xr - t) - te) * (y * y) * (1.0 / 3.0);
    const double a = xr * xr, ae = fma(xr, xr, -a);
    const double b = a * xr, be = fma(a, xr, -b) + ae * xr;
    const double hi = b * c0, lo = fma(b, c0, -hi) + (b * c1 + be * c0);
    const double u = 2.0 * f - 1.0;
    const double lnh = e * 0.6931471805599453 +
                       u * (0.9866915272798602 + u * (-0.40741277789409874 + u * 0.11478007077615415));
    const double d = hi * (1.4802973661668753e-16 * lnh);
    return ldexp(hi + (lo + d), 10 * q);
}

// Constants of one step's face fluxes, FP64 reference form.
struct FluxConstsF64 {
    double dx;      // face width and the Sf divisor
    double fn;      // Manning n
    double thresh;  // DepthThresh
    double maxh;    // MaxHflow
    double gdt;     // g*dt, evaluated first as the reference does
    double theta;
};

// Inertial flux through one interior face, the expression of
// src/kernels.cpp:60-89 (x faces) and :110-139 (y faces).
// z0/h0 are the upstream-index cell (i or j), z1/h1 the downstream one;
// q0 = q_old of the face, qup/qdown its neighbours along the flow axis,
// qt the transverse average of q_old.
__device__ __forceinline__ double face_flux_ref(double z0, double z1, double h0, double h1,
                                                double q0, double qup, double qdown, double qt,
                                                const FluxConstsF64& k) {
    double y0 = z0 + h0, y1 = z1 + h1;
    double dh = y0 - y1;
    double Sf = -dh / k.dx;
    double hflow = getmax(y0, y1) - getmax(z0, z1);
    hflow = getmax(hflow, 0.0);
    hflow = getmin(hflow, k.maxh);
    double Q = 0.0;
    if (hflow > k.thresh) {
        double qvect = sqrt(q0 * q0 + qt * qt);
        double num = (k.theta * q0 + 0.5 * (1.0 - k.theta) * (qup + qdown)) - (k.gdt * hflow * Sf);
        double den = 1.0 + k.gdt * hflow * k.fn * k.fn * fabs(qvect) / pow_10_3(hflow);
        Q = num / den * k.dx;
        if (Q * dh < 0.0) {
            num = q0 - (k.gdt * hflow * Sf);
            Q = num / den * k.dx;
        }
    }
    return Q;
}

// theta == 1 form of face_flux_ref.  With theta = 1 the reference's
// numerator is q0 + 0*(qup+qdown) - g*dt*hflow*Sf, equal in value to
// q0 - g*dt*hflow*Sf, and its sign-reversal branch recomputes that same
// value, so the result equals face_flux_ref's for every finite input.
__device__ __forceinline__ double face_flux_ref_t1(double z0, double z1, double h0, double h1,
                                                   double q0, double qt, const FluxConstsF64& k) {
    double y0 = z0 + h0, y1 = z1 + h1;
    double dh = y0 - y1;
    double hflow = getmax(y0, y1) - getmax(z0, z1);
    hflow = getmax(hflow, 0.0);
    hflow = getmin(hflow, k.maxh);
    double Q = 0.0;
    if (hflow > k.thresh) {
        double Sf = -dh / k.dx;
        double qvect = sqrt(q0 * q0 + qt * qt);
        double num = (k.theta * q0) - (k.gdt * hflow * Sf);
        double den = 1.0 + k.gdt * hflow * k.fn * k.fn * fabs(qvect) / pow_10_3(hflow);
        Q = num / den * k.dx;
    }
    return Q;
}

// face_flux_ref_t1 on stored discharges: Q0 is the face's Q and a..d the
// four transverse Q of the average; both become q_old = Q * sc (sc = 1/dx)
// exactly as UpdateQs_GPU scales them (kernels.cpp:8-16), only on wet faces.
__device__ __forceinline__ double face_flux_ref_t1_q(double z0, double z1, double h0, double h1,
                                                     double Q0, double a, double b, double c,
                                                     double d, double sc,
                                                     const FluxConstsF64& k) {
    double y0 = z0 + h0, y1 = z1 + h1;
    double dh = y0 - y1;
    double hflow = getmax(y0, y1) - getmax(z0, z1);
    hflow = getmax(hflow, 0.0);
    hflow = getmin(hflow, k.maxh);
    double Q = 0.0;
    if (hflow > k.thresh) {
        const double q0 = Q0 * sc;
        const double qt = (((a * sc + b * sc) + c * sc) + d * sc) / 4.0;
        double Sf = -dh / k.dx;
        double qvect = sqrt(q0 * q0 + qt * qt);
        double num = (k.theta * q0) - (k.gdt * hflow * Sf);
        double den = 1.0 + k.gdt * hflow * k.fn * k.fn * fabs(qvect) / pow_10_3(hflow);
        Q = num / den * k.dx;
    }
    return Q;
}

// Constants of one step's face fluxes, FP32 form.
struct FluxConstsF32 {
    float dx;
    float thresh;
    float maxh;
    float gdt_dx;   // g*dt/dx: num = q0 + (g*dt/dx)*hflow*dh
    float gdt_n2;   // g*dt*n^2: den = 1 + g*dt*n^2*|q|*hflow^(-7/3)
    float sc;       // 1/dx: Q -> q_old
    float sc4;      // 0.25/dx: sum of four Q -> transverse q_old average
    float zsc;      // metres per unit of the fixed-point terrain (a power of two)
};

// Terrain step z0 - z1 of a face from the FP32 path's fixed-point terrain
// (integer units of zsc metres).  The integer difference is exact (taken
// modulo 2^32, which no valid pair reaches; a NODATA pair's value is never
// used), so dz carries one FP32 rounding of the step instead of the
// rounding of each elevation: stored as float, an elevation 50 m above the
// lowest cell is known to about 4e-6 m, as fixed point to zsc/2 (6e-8 m for
// 80 m of relief).
__device__ __forceinline__ float terrain_step(int z0, int z1, float zsc) {
    return (float)(int)((unsigned)z0 - (unsigned)z1) * zsc;
}

// Surface step and flow depth of a face in FP32 without forming the surface
// elevations z + h: tens of metres of terrain round to about 1e-6 m in FP32,
// while the terrain step dz = z0 - z1 and the depths are small numbers.
//   dh    = (z0 + h0) - (z1 + h1)          = dz + (h0 - h1)
//   hflow = max(z0+h0, z1+h1) - max(z0, z1) = dz >= 0 ? max(h0, h1 - dz)
//                                                     : max(h0 + dz, h1)
__device__ __forceinline__ void face_depths_f32(float dz, float h0, float h1, float& dh,
                                                float& hflow) {
    dh = dz + (h0 - h1);
    hflow = (dz >= 0.0f) ? getmax(h0, h1 - dz) : getmax(h0 + dz, h1);
}

// FP32 form of face_flux_ref_t1.  Same physics, rearranged for the VALU:
//   g*dt*hflow*Sf          = -(g*dt/dx)*hflow*dh
//   g*dt*hflow*n^2*|q| / hflow^(10/3) = g*dt*n^2*|q| * hflow^(-7/3)
// with hflow^(-7/3) = exp2(-7/3*log2(hflow)) on the hardware transcendental
// units and one reciprocal for num/den.  The wet-face formula is evaluated on
// every lane and the result selected by hflow > DepthThresh: a divergent
// branch per face costs more exec-mask instructions than the formula, and a
// lone wavefront's row is bound by the instructions it issues.  (On a dry
// lane the discarded value may be Inf or NaN; it is never used.)
__device__ __forceinline__ float face_flux_f32(float dz, float h0, float h1, float q0, float qt,
                                               const FluxConstsF32& k) {
    float dh, hflow;
    face_depths_f32(dz, h0, h1, dh, hflow);
    hflow = getmax(hflow, 0.0f);
    hflow = getmin(hflow, k.maxh);
    const float qvect = __builtin_amdgcn_sqrtf(q0 * q0 + qt * qt);
    const float num = q0 + k.gdt_dx * hflow * dh;
    const float p = __builtin_amdgcn_exp2f(-2.33333333f * __builtin_amdgcn_logf(hflow));
    const float den = 1.0f + k.gdt_n2 * qvect * p;
    const float Q = num * k.dx * __builtin_amdgcn_rcpf(den);
    return (hflow > k.thresh) ? Q : 0.0f;
}

// face_flux_f32 on stored discharges (see face_flux_ref_t1_q).
__device__ __forceinline__ float face_flux_f32_q(float dz, float h0, float h1, float Q0, float a,
                                                 float b, float c, float d,
                                                 const FluxConstsF32& k) {
    const float q0 = Q0 * k.sc;
    const float qt = (((a + b) + c) + d) * k.sc4;
    return face_flux_f32(dz, h0, h1, q0, qt, k);
}

// Edge fluxes of src/kernels.cpp:154-253 from the pre-update depth.
// West (kernels.cpp:155-179): needs the inner neighbour (z1, h1) unchecked.
template <typename T>
__device__ __forceinline__ T edge_west(T z0, T h0, T z1, T h1, bool nodata0, T dx, T g) {
    if (nodata0 || h0 < T(0.001)) return T(0);
    T y0 = z0 + h0, y1 = z1 + h1;
    if (y0 > z0) {
        T Sf = (y1 - y0) / dx;
        if (Sf < 0) return -h0 * sqrt(g * h0) * dx * T(0.5);
        return T(0);
    }
    return T(0);
}
// edge_west in FP32 from the terrain step dz10 = z1 - z0: the surface step
// y1 - y0 = dz10 + (h1 - h0), and y0 > z0 holds for every h0 >= 0.001 (the
// elevation is finite), so the test reduces to h0 > 0.
__device__ __forceinline__ float edge_west_f32(float h0, float h1, float dz10, bool nodata0,
                                               float dx, float g) {
    if (nodata0 || h0 < 0.001f) return 0.0f;
    if (h0 > 0.0f) {
        const float Sf = (dz10 + (h1 - h0)) / dx;
        if (Sf < 0.0f) return -h0 * sqrtf(g * h0) * dx * 0.5f;
        return 0.0f;
    }
    return 0.0f;
}

// East and South (kernels.cpp:188-202, 234-248): outflow +h0*sqrt(g*h0)*dx/2.
template <typename T>
__device__ __forceinline__ T edge_out_pos(T h0, bool nodata0, T dx, T g) {
    if (nodata0 || h0 < T(0.001)) return T(0);
    if (h0 > T(0.001)) return h0 * sqrt(g * h0) * dx * T(0.5);
    return T(0);
}
// North (kernels.cpp:211-225): outflow -h0*sqrt(g*h0)*dx/2.
template <typename T>
__device__ __forceinline__ T edge_out_neg(T h0, bool nodata0, T dx, T g) {
    if (nodata0 || h0 < T(0.001)) return T(0);
    if (h0 > T(0.001)) return -h0 * sqrt(g * h0) * dx * T(0.5);
    return T(0);
}

// Order-preserving maps from IEEE values to unsigned keys, so that an
// unsigned atomicMax computes a floating-point max (NaN excluded by callers).
__host__ __device__ __forceinline__ unsigned long long fkey(double x) {
    unsigned long long b;
    __builtin_memcpy(&b, &x, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double fkey_inv_d(unsigned long long k) {
    unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}
__host__ __device__ __forceinline__ unsigned int fkey(float x) {
    unsigned int b;
    __builtin_memcpy(&b, &x, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ __forceinline__ float fkey_inv_f(unsigned int k) {
    unsigned int b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

// Linear interpolation of src/io.cpp:299-311: clamped at both ends, bracket
// [t_i, t_{i+1}), empty series -> empty_value.
template <typename PD>
__host__ __device__ __forceinline__ double interp_series(PD times, PD values, int n,
                                                         double empty_value, double t) {
    if (n <= 0) return empty_value;
    if (t <= times[0]) return values[0];
    if (t >= times[n - 1]) return values[n - 1];
    for (int i = 0; i < n - 1; i++) {
        if (t >= times[i] && t < times[i + 1]) {
            double ratio = (t - times[i]) / (times[i + 1] - times[i]);
            return values[i] + ratio * (values[i + 1] - values[i]);
        }
    }
    return values[n - 1];
}

// interp_series with every load of a short series (up to four points, as a
// hydrograph or a rain record usually is) issued at once, so that a value
// costs one round trip to memory instead of one per bracket visited.  The
// indices are clamped to the last point, which makes the padded brackets
// empty; the end tests and the brackets are visited in interp_series' order
// with its arithmetic, so the value is the same for every input.
template <typename PD>
__host__ __device__ __forceinline__ double interp_series_short(PD times, PD values, int n,
                                                      double empty_value, double t) {
    if (n <= 0) return empty_value;
    if (n > 4) return interp_series(times, values, n, empty_value, t);
    double tt[4], vv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int c = (i < n) ? i : n - 1;
        tt[i] = times[c];
        vv[i] = values[c];
    }
    if (t <= tt[0]) return vv[0];
    if (t >= tt[3]) return vv[3];  // tt[3] is times[n-1]
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (t >= tt[i] && t < tt[i + 1]) {
            double ratio = (t - tt[i]) / (tt[i + 1] - tt[i]);
            return vv[i] + ratio * (vv[i + 1] - vv[i]);
        }
    }
    return vv[3];
}

// dt of src/kernels.cpp:36-43
__host__ __device__ __forceinline__ double cfl_dt(double hmax, double thresh, double cfl,
                                                  double dx, double g, double init_tstep) {
    double dt;
    if (hmax > thresh) dt = cfl * dx / sqrt(g * hmax);
    else dt = init_tstep;
    return dt;
}

}  // namespace af
