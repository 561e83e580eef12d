/*
 * oracle.c -- FP64 CPU oracle for the AdaptFlood local-inertial time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (adaptflood_amd/, include/)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg do, and only as the checker / the CPU baseline.
 *
 * It restates, in plain C, the arithmetic of the reference's unfused kernel
 * sequence in the reference's own operation order (compile with
 * -ffp-contract=off so no FMA is formed that the reference would not form):
 *
 *   orc_update_qs          src/kernels.cpp:6-19     (UpdateQs_GPU)
 *   orc_find_hmax          src/kernels.cpp:28-33    (FindHmax reduce)
 *   orc_floodplain_q       src/kernels.cpp:22-146   (FloodplainQ_GPU: dt, Qx, Qy)
 *   orc_bcs                src/kernels.cpp:149-256  (BCs_GPU)
 *   orc_apply_point_bcs    src/kernels.cpp:259-277  (ApplyPointBCs_GPU)
 *   orc_update_h           src/kernels.cpp:280-290  (UpdateH_GPU)
 *   orc_rainfall           src/kernels.cpp:354-367  (Rainfall_GPU)
 *   orc_add_flow_bcs       src/main.cpp:256-260     (inline AddFlowBCs kernel)
 *   orc_interp             src/io.cpp:299-311 and io.cpp:88-100 (interpolation)
 *   orc_run                src/main.cpp:196-304     (time loop + mass report)
 *   orc_velocity           src/kernels.cpp:293-351  (ComputeVelocity_GPU)
 *   orc_pow_10_3           src/kernels.cpp:83, 133  (pow(hflow, 10.0/3.0))
 *
 * Layout (reference): cells idx = i + j*xsz (row 0 = north); face arrays Qx,
 * Qy, Qxold, Qyold have (xsz+1)*(ysz+1) entries with pitch xsz+1.
 *
 * Point boundaries with duplicate cells: the reference's AddFlowBCs and
 * ApplyPointBCs kernels race on duplicates (non-atomic read-modify-write,
 * last writer wins).  The oracle applies them sequentially in file order,
 * which is the one deterministic outcome the reference can produce.
 *
 * Parity pins: tests/test_oracle.py checks each function here against hand
 * derived known answers, closed forms and invariants (dt closed form, a
 * single-face flux evaluated independently in long double, per-edge BC closed
 * forms, source volume v*dx*dt, rain volume r*t*A_valid on a closed basin,
 * lake-at-rest, closed-domain mass conservation, mirror symmetry, Manning
 * normal depth on a plane, a hand-computed velocity field, and orc_pow_10_3
 * against the C library's pow and against 80-digit decimal values);
 * tests/test_golden.py holds the oracle to the committed trajectory of the
 * 256 x 256 plane workload.  The reference itself cannot be built in this
 * environment (it needs Kokkos), so no vector from a reference run exists.
 *
 * pow: the reference takes pow(hflow, 10.0/3.0) from its backend's libm
 * (kernels.cpp:83, :133).  The oracle uses orc_pow_10_3, a pow built from
 * IEEE-exact operations that is within 1 ulp of glibc's everywhere, so that
 * a device implementation of the same operation sequence agrees with the
 * oracle bit for bit; see orc_pow_10_3 below.
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Clamp accounting.  Every clamp of a depth at zero (kernels.cpp:287,
 * kernels.cpp:363, main.cpp:259) adds the depth it removes to the water.
 * Built with -DORC_COUNT_CLAMP (the third library of oracle/Makefile), the
 * oracle sums those depths, so that a run's volume budget closes:
 *   V(t) - V(0) = sources + rain + edge inflow + clamp gains + overrides.
 * The default build clamps the same way and counts nothing. */
#ifdef ORC_COUNT_CLAMP
static double orc_clamp_total = 0.0;
#define ORC_CLAMP_AT_ZERO(h, acc)  \
    do {                           \
        if ((h) < 0.0) {           \
            (acc) -= (h);          \
            (h) = 0.0;             \
        }                          \
    } while (0)
#define ORC_GAIN_REDUCTION reduction(+ : gain)
#define ORC_GAIN_ADD(acc) (orc_clamp_total += (acc))
#else
#define ORC_CLAMP_AT_ZERO(h, acc)  \
    do {                           \
        if ((h) < 0.0) (h) = 0.0;  \
    } while (0)
#define ORC_GAIN_REDUCTION
#define ORC_GAIN_ADD(acc) ((void)(acc))
#endif

/* The depth the clamps have added since the last reset (a sum of depths;
 * times dA for a volume), or -1 in a build that does not count. */
double orc_clamp_gain(int reset) {
#ifdef ORC_COUNT_CLAMP
    double g = orc_clamp_total;
    if (reset) orc_clamp_total = 0.0;
    return g;
#else
    (void)reset;
    return -1.0;
#endif
}

static inline double getmax(double a, double b) { return (a > b) ? a : b; } /* types.hpp:14 */
static inline double getmin(double a, double b) { return (a < b) ? a : b; } /* types.hpp:15 */

/* pow(hflow, 10.0/3.0) of kernels.cpp:83 and :133.  The reference takes it
 * from its backend's libm, whose last bit differs between libraries (the
 * CPU build's glibc and the GPU build's device library disagree in a
 * fraction of a percent of arguments).  The oracle fixes one rounding: a
 * pow built from IEEE-exact operations only, so a device implementation of
 * the same operation sequence reproduces it bit for bit.
 *   h = xr 2^(3q), xr in [1, 8); h^(10/3) = xr^3 cbrt(xr) 2^(10q)
 *   cbrt: Newton on the inverse cube root (y <- y(4 - xr y^3)/3, 6 steps
 *   from a quadratic start), then one Newton correction with the residual
 *   in double-double (Dekker products through fma); xr^3 in double-double;
 *   the exponent's excess 10.0/3.0 - 10/3 = 1.4802973661668753e-16 enters
 *   as the first-order factor 1 + 1.48e-16 ln h.
 * Pinned by tests/test_oracle.py against the C library's pow (at most 1 ulp
 * apart) and against 80-digit decimal values (correctly rounded in all but
 * a few in 10^4). */
double orc_pow_10_3(double h) {
    int E;
    double f = frexp(h, &E);
    int e = E - 1;
    int q = (e >= 0) ? e / 3 : -((2 - e) / 3);
    double xr = ldexp(2.0 * f, e - 3 * q);
    double y = 1.190645532750796 + xr * (-0.20554765291607405 + xr * 0.01490212016527807);
    for (int i = 0; i < 6; i++) y = y * (4.0 - xr * (y * y * y)) * (1.0 / 3.0);
    double c0 = xr * (y * y);
    double p = c0 * c0, pe = fma(c0, c0, -p);
    double t = p * c0, te = fma(p, c0, -t) + pe * c0;
    double c1 = ((xr - t) - te) * (y * y) * (1.0 / 3.0);
    double a = xr * xr, ae = fma(xr, xr, -a);
    double b = a * xr, be = fma(a, xr, -b) + ae * xr;
    double hi = b * c0, lo = fma(b, c0, -hi) + (b * c1 + be * c0);
    double u = 2.0 * f - 1.0;
    double lnh = e * 0.6931471805599453 +
                 u * (0.9866915272798602 + u * (-0.40741277789409874 + u * 0.11478007077615415));
    double d = hi * (1.4802973661668753e-16 * lnh);
    return ldexp(hi + (lo + d), 10 * q);
}

/* The flux's pow(hflow, 10.0/3.0).  Built with -DORC_LIBM_POW (the second
 * library of oracle/Makefile), it is the C library's pow, as in the
 * reference's CPU build (kernels.cpp:83, 133): that build measures how far
 * the choice of pow rounding moves a whole run from the pinned oracle. */
#ifdef ORC_LIBM_POW
#define ORC_POW_10_3(h) pow((h), 10.0 / 3.0)
#else
#define ORC_POW_10_3(h) orc_pow_10_3(h)
#endif

/* kernels.cpp:6-19 */
void orc_update_qs(const double* Qx, const double* Qy, double* Qxold, double* Qyold,
                   int xsz, int ysz, double dx) {
    double dxinv = 1.0 / dx;
    long n = (long)(xsz + 1) * (long)(ysz + 1);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) Qxold[k] = Qx[k] * dxinv;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) Qyold[k] = Qy[k] * dxinv;
}

/* kernels.cpp:28-33: max of H over cells whose DEM is not nodata.  The
 * reduction identity of a max reducer is the lowest double. */
double orc_find_hmax(const double* H, const double* DEM, long n, double nodata) {
    double m = -DBL_MAX;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long k = 0; k < n; k++) {
        if (DEM[k] != nodata && H[k] > m) m = H[k];
    }
    return m;
}

/* kernels.cpp:36-43 */
double orc_dt_from_hmax(double Hmax, double DepthThresh, double cfl, double dx, double g,
                        double InitTstep) {
    double dt;
    if (Hmax > DepthThresh) dt = cfl * dx / sqrt(g * Hmax);
    else dt = InitTstep;
    return dt;
}

/* One interior face of kernels.cpp:58-92 (x) / 108-142 (y).  q0 is the face's
 * own q_old, qup/qdown the neighbours along the flow direction, qt the
 * transverse average.  Returns Q in m3/s. */
double orc_face_flux(double z0, double z1, double h0, double h1, double q0, double qup,
                     double qdown, double qt, double dx, double fn, double DepthThresh,
                     double MaxHflow, double dt, double g, double theta) {
    double y0 = z0 + h0, y1 = z1 + h1;
    double dh = y0 - y1;
    double Sf = -dh / dx;
    double hflow = getmax(y0, y1) - getmax(z0, z1);
    hflow = getmax(hflow, 0.0);
    hflow = getmin(hflow, MaxHflow);
    double Q = 0.0;
    if (hflow > DepthThresh) {
        double qvect = sqrt(q0 * q0 + qt * qt);
        double numerator = (theta * q0 + 0.5 * (1.0 - theta) * (qup + qdown)) - (g * dt * hflow * Sf);
        double denominator = 1.0 + g * dt * hflow * fn * fn * fabs(qvect) / ORC_POW_10_3(hflow);
        Q = numerator / denominator * dx;
        if (Q * dh < 0.0) {
            numerator = q0 - (g * dt * hflow * Sf);
            Q = numerator / denominator * dx;
        }
    }
    return Q;
}

/* kernels.cpp:22-146.  Tstep is an out-parameter. */
void orc_floodplain_q(const double* H, const double* DEM, double* Qx, double* Qy,
                      const double* Qxold, const double* Qyold, int xsz, int ysz, double dx,
                      double dy, double FPn, double DepthThresh, double MaxHflow, double dhlin,
                      double* Tstep, double Qlimfact, double dA, double cfl, double g,
                      double InitTstep, double theta, double nodata) {
    (void)dy; (void)dhlin; (void)Qlimfact; (void)dA;
    double Hmax = orc_find_hmax(H, DEM, (long)xsz * ysz, nodata);
    double dt = orc_dt_from_hmax(Hmax, DepthThresh, cfl, dx, g, InitTstep);
    *Tstep = dt;
    const long P = xsz + 1;
    /* x faces: i in [0, xsz-2], j in [0, ysz-1] -> Qx[(i+1) + j*P] */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ysz; j++) {
        for (int i = 0; i < xsz - 1; i++) {
            long p0 = i + (long)j * xsz, p1 = p0 + 1;
            long pq0 = (i + 1) + (long)j * P;
            double z0 = DEM[p0], z1 = DEM[p1];
            if (z0 == nodata || z1 == nodata) { Qx[pq0] = 0.0; continue; }
            double q0 = Qxold[pq0];
            double qup = (i > 0) ? Qxold[pq0 - 1] : 0.0;
            double qdown = (i < xsz - 2) ? Qxold[pq0 + 1] : 0.0;
            double qt = (Qyold[i + (long)j * P] + Qyold[(i + 1) + (long)j * P] +
                         Qyold[i + (long)(j + 1) * P] + Qyold[(i + 1) + (long)(j + 1) * P]) / 4.0;
            Qx[pq0] = orc_face_flux(z0, z1, H[p0], H[p1], q0, qup, qdown, qt, dx, FPn,
                                    DepthThresh, MaxHflow, dt, g, theta);
        }
    }
    /* y faces: i in [0, xsz-1], j in [0, ysz-2] -> Qy[i + (j+1)*P] */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ysz - 1; j++) {
        for (int i = 0; i < xsz; i++) {
            long p0 = i + (long)j * xsz, p1 = i + (long)(j + 1) * xsz;
            long pq0 = i + (long)(j + 1) * P;
            double z0 = DEM[p0], z1 = DEM[p1];
            if (z0 == nodata || z1 == nodata) { Qy[pq0] = 0.0; continue; }
            double q0 = Qyold[pq0];
            double qup = (j > 0) ? Qyold[i + (long)j * P] : 0.0;
            double qdown = (j < ysz - 2) ? Qyold[i + (long)(j + 2) * P] : 0.0;
            double qt = (Qxold[i + (long)j * P] + Qxold[(i + 1) + (long)j * P] +
                         Qxold[i + (long)(j + 1) * P] + Qxold[(i + 1) + (long)(j + 1) * P]) / 4.0;
            Qy[pq0] = orc_face_flux(z0, z1, H[p0], H[p1], q0, qup, qdown, qt, dx, FPn,
                                    DepthThresh, MaxHflow, dt, g, theta);
        }
    }
}

/* kernels.cpp:149-256.  H is the pre-update depth; dt is unused. */
void orc_bcs(double* Qx, double* Qy, const double* H, const double* DEM, int xsz, int ysz,
             double dx, double g, double dt, double nodata, int west_free, int east_free,
             int north_free, int south_free) {
    (void)dt;
    const long P = xsz + 1;
    if (west_free) {
        for (int j = 0; j < ysz; j++) {
            long p0 = (long)j * xsz;
            double h0 = H[p0], z0 = DEM[p0];
            double q = 0.0;
            if (!(z0 == nodata || h0 < 0.001)) {
                long p1 = 1 + (long)j * xsz;
                double y0 = z0 + h0, y1 = DEM[p1] + H[p1];
                if (y0 > z0) {
                    double Sf = (y1 - y0) / dx;
                    if (Sf < 0) q = -h0 * sqrt(g * h0) * dx * 0.5;
                }
            }
            Qx[(long)j * P] = q;
        }
    } else {
        for (int j = 0; j < ysz + 1; j++) Qx[(long)j * P] = 0.0;
    }
    if (east_free) {
        for (int j = 0; j < ysz; j++) {
            long p0 = (xsz - 1) + (long)j * xsz;
            double h0 = H[p0], z0 = DEM[p0];
            double q = 0.0;
            if (!(z0 == nodata || h0 < 0.001)) {
                if (h0 > 0.001) q = h0 * sqrt(g * h0) * dx * 0.5;
            }
            Qx[xsz + (long)j * P] = q;
        }
    } else {
        for (int j = 0; j < ysz + 1; j++) Qx[xsz + (long)j * P] = 0.0;
    }
    if (north_free) {
        for (int i = 0; i < xsz; i++) {
            double h0 = H[i], z0 = DEM[i];
            double q = 0.0;
            if (!(z0 == nodata || h0 < 0.001)) {
                if (h0 > 0.001) q = -h0 * sqrt(g * h0) * dx * 0.5;
            }
            Qy[i] = q;
        }
    } else {
        for (int i = 0; i < xsz + 1; i++) Qy[i] = 0.0;
    }
    if (south_free) {
        for (int i = 0; i < xsz; i++) {
            long p0 = i + (long)(ysz - 1) * xsz;
            double h0 = H[p0], z0 = DEM[p0];
            double q = 0.0;
            if (!(z0 == nodata || h0 < 0.001)) {
                if (h0 > 0.001) q = h0 * sqrt(g * h0) * dx * 0.5;
            }
            Qy[i + (long)ysz * P] = q;
        }
    } else {
        for (int i = 0; i < xsz + 1; i++) Qy[i + (long)ysz * P] = 0.0;
    }
}

/* kernels.cpp:259-277, applied in file order (last writer wins). */
void orc_apply_point_bcs(double* H, const int* bc_idx, const int* bc_type,
                         const double* bc_value, int num_bcs) {
    for (int b = 0; b < num_bcs; b++) {
        int type = bc_type[b];
        if (type == 1) H[bc_idx[b]] = 0.0;
        else if (type == 2 || type == 3) H[bc_idx[b]] = bc_value[b];
    }
}

/* kernels.cpp:280-290 */
void orc_update_h(double* H, const double* Qx, const double* Qy, int xsz, int ysz, double dA,
                  double Tstep) {
    const long P = xsz + 1;
    double gain = 0.0;
#pragma omp parallel for schedule(static) ORC_GAIN_REDUCTION
    for (int j = 0; j < ysz; j++) {
        for (int i = 0; i < xsz; i++) {
            long idx = i + (long)j * xsz;
            double Qx_in = Qx[i + (long)j * P], Qx_out = Qx[(i + 1) + (long)j * P];
            double Qy_in = Qy[i + (long)j * P], Qy_out = Qy[i + (long)(j + 1) * P];
            H[idx] += Tstep * (Qx_in - Qx_out + Qy_in - Qy_out) / dA;
            ORC_CLAMP_AT_ZERO(H[idx], gain);
        }
    }
    ORC_GAIN_ADD(gain);
}

/* kernels.cpp:354-367 */
void orc_rainfall(double* H, const double* DEM, int xsz, int ysz, double rain_rate, double Tstep,
                  double nodata) {
    if (rain_rate <= 0.0) return;
    double rain_depth = rain_rate * Tstep;
    long n = (long)xsz * ysz;
    double gain = 0.0;
#pragma omp parallel for schedule(static) ORC_GAIN_REDUCTION
    for (long k = 0; k < n; k++) {
        if (DEM[k] != nodata) {
            H[k] += rain_depth;
            ORC_CLAMP_AT_ZERO(H[k], gain);
        }
    }
    ORC_GAIN_ADD(gain);
}

/* main.cpp:256-260, applied in file order. */
void orc_add_flow_bcs(double* H, const int* ps_idx, const double* ps_flow, int n) {
    double gain = 0.0;
    for (int s = 0; s < n; s++) {
        int idx = ps_idx[s];
        H[idx] += ps_flow[s];
        ORC_CLAMP_AT_ZERO(H[idx], gain);
    }
    ORC_GAIN_ADD(gain);
}

/* io.cpp:299-311 (interpolate_bc_value) and io.cpp:88-100 (interpolate_rain):
 * linear, clamped at both ends, bracket [t_i, t_{i+1}).  An empty series
 * returns `empty_value` (the point's fixed value, or 0 for rain). */
double orc_interp(const double* times, const double* values, int n, double empty_value,
                  double t) {
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

/* kernels.cpp:293-351 */
void orc_velocity(double* Vx, double* Vy, const double* H, const double* DEM, const double* Qx,
                  const double* Qy, int xsz, int ysz, double dx, double nodata) {
    const long P = xsz + 1;
    for (int j = 0; j < ysz; j++) {
        for (int i = 0; i < xsz; i++) {
            long idx = i + (long)j * xsz;
            double z = DEM[idx], h = H[idx];
            if (z == nodata || h < 0.001) { Vx[idx] = 0.0; Vy[idx] = 0.0; continue; }
            double hflow = h;
            if (hflow < 0.001) hflow = 0.001;
            double a = Qx[i + (long)j * P], b = Qx[(i + 1) + (long)j * P];
            double va = (fabs(a) > 1e-10) ? a / dx / hflow : 0.0;
            double vb = (fabs(b) > 1e-10) ? b / dx / hflow : 0.0;
            Vx[idx] = (fabs(va) > fabs(vb)) ? va : vb;
            double c = Qy[i + (long)j * P], d = Qy[i + (long)(j + 1) * P];
            double vc = (fabs(c) > 1e-10) ? c / dx / hflow : 0.0;
            double vd = (fabs(d) > 1e-10) ? d / dx / hflow : 0.0;
            Vy[idx] = (fabs(vc) > fabs(vd)) ? vc : vd;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Whole-run driver: main.cpp:196-304 without file output.                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    /* grid */
    int xsz, ysz;
    double dx, nodata;
    /* physics, main.cpp:115-127 */
    double fpn, depth_thresh, max_hflow, init_tstep, cfl, g, theta;
    double sim_time, mass_int;
    /* line BCs, main.cpp:62-70 */
    int west_free, east_free, north_free, south_free;
    /* flow points (QFIX=4, QVAR=5), in file order */
    int n_flow;
    const int* flow_idx;
    const int* flow_type;
    const double* flow_value;
    const int* flow_off; /* series offset into series_t / series_v */
    const int* flow_len;
    /* other points (FREE=1, HFIX=2, HVAR=3 as bc_type codes), in file order */
    int n_other;
    const int* other_idx;
    const int* other_code;
    const double* other_value;
    const int* other_off;
    const int* other_len;
    const double* series_t;
    const double* series_v;
    /* rain (m/s); rain_enabled as main.cpp:265 */
    int rain_enabled;
    int n_rain;
    const double* rain_t;
    const double* rain_r;
} orc_case;

typedef struct {
    long max_steps;   /* stop after this many steps even if t < sim_time (<=0: no limit) */
    long steps_done;
    double t;
    double* dt_log;   /* optional, length dt_log_cap */
    long dt_log_cap;
    double* mass_log; /* optional, 4 doubles per report: t, Tstep, Vol, Area */
    long mass_log_cap;
    long n_mass;
    double step_seconds; /* wall time of the step loop (excluding mass reports) */
    double next_mass;    /* next mass-report time after the run */
} orc_run_info;

/* Runs the time loop of main.cpp:227-304 on caller-owned state.  H and DEM are
 * xsz*ysz; Qx, Qy, Qxold, Qyold are (xsz+1)*(ysz+1).  The caller zero-fills the
 * state as main.cpp:156-160 does (or passes a state to continue from); t0 and
 * next_mass carry the clock between calls. */
int orc_run(const orc_case* c, double* H, const double* DEM, double* Qx, double* Qy,
            double* Qxold, double* Qyold, double t0, double next_mass0, orc_run_info* info) {
    const int xsz = c->xsz, ysz = c->ysz;
    const double dx = c->dx, dA = dx * dx;
    double* ps_flow = (double*)malloc(sizeof(double) * (c->n_flow > 0 ? c->n_flow : 1));
    double* bc_value = (double*)malloc(sizeof(double) * (c->n_other > 0 ? c->n_other : 1));
    if (!ps_flow || !bc_value) { free(ps_flow); free(bc_value); return -1; }
    for (int b = 0; b < c->n_other; b++) bc_value[b] = c->other_value[b];
    double t = t0, next_mass = next_mass0;
    long step = 0;
    info->n_mass = 0;
    double loop_s = 0.0;
    while (t < c->sim_time) {
        if (info->max_steps > 0 && step >= info->max_steps) break;
        struct timespec ta, tb;
        clock_gettime(CLOCK_MONOTONIC, &ta);
        double Tstep = c->init_tstep; /* main.cpp:228 */
        orc_floodplain_q(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, dx, c->fpn, c->depth_thresh,
                         c->max_hflow, 0.01, &Tstep, 1.0, dA, c->cfl, c->g, c->init_tstep,
                         c->theta, c->nodata);
        orc_bcs(Qx, Qy, H, DEM, xsz, ysz, dx, c->g, Tstep, c->nodata, c->west_free, c->east_free,
                c->north_free, c->south_free);
        if (c->n_flow > 0) { /* main.cpp:245-261 */
            for (int s = 0; s < c->n_flow; s++) {
                double v;
                if (c->flow_type[s] == 4) v = c->flow_value[s];
                else v = orc_interp(c->series_t + c->flow_off[s], c->series_v + c->flow_off[s],
                                    c->flow_len[s], c->flow_value[s], t);
                ps_flow[s] = v * dx * Tstep / dA;
            }
            orc_add_flow_bcs(H, c->flow_idx, ps_flow, c->n_flow);
        }
        if (c->rain_enabled) { /* main.cpp:265-268 */
            double r = orc_interp(c->rain_t, c->rain_r, c->n_rain, 0.0, t);
            orc_rainfall(H, DEM, xsz, ysz, r, Tstep, c->nodata);
        }
        orc_update_h(H, Qx, Qy, xsz, ysz, dA, Tstep);
        if (c->n_other > 0) { /* main.cpp:274-282 */
            for (int b = 0; b < c->n_other; b++) {
                if (c->other_code[b] == 3)
                    bc_value[b] = orc_interp(c->series_t + c->other_off[b],
                                             c->series_v + c->other_off[b], c->other_len[b],
                                             c->other_value[b], t);
            }
            orc_apply_point_bcs(H, c->other_idx, c->other_code, bc_value, c->n_other);
        }
        orc_update_qs(Qx, Qy, Qxold, Qyold, xsz, ysz, dx);
        clock_gettime(CLOCK_MONOTONIC, &tb);
        loop_s += (double)(tb.tv_sec - ta.tv_sec) + 1e-9 * (double)(tb.tv_nsec - ta.tv_nsec);
        if (info->dt_log && step < info->dt_log_cap) info->dt_log[step] = Tstep;
        t += Tstep;
        step++;
        if (t >= next_mass) { /* main.cpp:291-304, summed in index order */
            double vol = 0.0, area = 0.0;
            long n = (long)xsz * ysz;
            for (long k = 0; k < n; k++) vol += H[k] * dA;
            for (long k = 0; k < n; k++) if (H[k] > 0.01) area += dA;
            if (info->mass_log && info->n_mass < info->mass_log_cap) {
                double* r = info->mass_log + 4 * info->n_mass;
                r[0] = t; r[1] = Tstep; r[2] = vol; r[3] = area;
            }
            info->n_mass++;
            next_mass += c->mass_int;
        }
    }
    info->steps_done = step;
    info->t = t;
    info->step_seconds = loop_s;
    info->next_mass = next_mass;
    free(ps_flow);
    free(bc_value);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
