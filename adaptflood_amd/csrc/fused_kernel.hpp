// fused_kernel.hpp -- one HIP kernel per local-inertial step.
//
// One launch advances every scenario of a batch by one step of
// src/main.cpp:227-288: CFL dt from the previous step's H maximum, x and y
// face fluxes (kernels.cpp:46-143), edge faces (kernels.cpp:149-256), point
// sources (main.cpp:244-262), rain (kernels.cpp:354-367), the depth update
// (kernels.cpp:280-290), FREE/HFIX/HVAR overrides (kernels.cpp:259-277) and
// the next step's H maximum, in that order per cell.  q_old is never stored:
// the kernel keeps Q and multiplies by 1/dx on read, which is the value
// UpdateQs_GPU (kernels.cpp:6-19) would have stored.
//
// Work decomposition (CDNA wave64): each wavefront owns a tile of one
// scenario, a strip of W = 64*V columns by up to R rows, and marches down the
// rows.  Lane l holds V adjacent columns, loaded as one 16-byte vector per
// array per row.  Neighbour values across lanes move with cross-lane shuffles;
// the two strip-edge columns come from one extra load per array per row.  A
// y face is computed once and carried to the next row in registers; an x face
// is computed once, except the strip's east face, which the last lane also
// computes (its neighbour strip computes it too; the arithmetic is identical).
// State ping-pongs between two buffers; the per-scenario control block carries
// t (FP64), the H maximum (three rotating slots, max-reduced with one atomic
// per wavefront) and the mass-report bookkeeping.
#pragma once
#include <float.h>
#include <limits.h>

#include "af_device.hpp"

namespace af {

// Per-scenario control block, device resident.
struct Ctl {
    double t[2];                  // t at the start of step k lives in t[k & 1]
    double next_mass[2];          // main.cpp:200,291,303, same parity as t
    unsigned long long hmax[3];   // fkey(H max) of the state entering step k: hmax[k % 3]
    long long n_done;             // executed steps
    double last_dt;
    int pending;                  // a mass report's partial sums are waiting
    int pending_buf;              // which partial buffer holds them
    double pending_t, pending_dt;
    int n_reports;
    int pad_;
};

// Per-scenario constants and device pointers.
struct ScnDev {
    int nx, ny;
    long long P;                  // row pitch in elements (all state arrays)
    int nstrips, tile_rows;
    int tile_base;                // first global tile of this scenario
    int ntiles;
    void* z;                      // T[ny*P]; NODATA cells hold `nodata_t`
    void* H[2];                   // T[ny*P]
    void* Qx[2];                  // T[ny*P], face f of row j at j*P + f, f in [0, nx]
    void* Qy[2];                  // T[(ny+1)*P], face row g (north of row g) at g*P + i
    double dx, dxinv, dA, fn, thresh, maxh, init_tstep, cfl, g, theta, sim_time, mass_int;
    double nodata_t;              // sentinel compared against z (as T)
    int west_free, east_free, north_free, south_free;
    // point sources (QFIX/QVAR) and overrides (FREE/HFIX/HVAR), sorted per
    // tile by (row, column, file order); *_begin has ntiles+1 entries
    const int* src_begin;
    const int* src_j;
    const int* src_i;
    const int* src_pt;
    const int* ovr_begin;
    const int* ovr_j;
    const int* ovr_i;
    const int* ovr_pt;
    const int* pt_type;
    const double* pt_value;
    const int* pt_off;
    const int* pt_len;
    const double* series_t;
    const double* series_v;
    int rain_enabled, n_rain;
    const double* rain_t;
    const double* rain_r;
    Ctl* ctl;
    double* part_vol;             // [2][ntiles]
    double* part_comp;            // [2][ntiles]
    unsigned long long* part_area;// [2][ntiles] wet-cell counts
    double* dt_log;
    long long dt_log_cap;
    double* reports;              // [report_cap][4]
    int report_cap;
    int pad_;
};

// Tile descriptor: scenario, strip, segment, and 1 for the scenario's first
// tile (whose wavefront also does the scenario's bookkeeping).
struct TileDesc {
    int scn, strip, seg, first;
};

template <typename T, int V>
struct alignas(sizeof(T) * V) Pack {
    T v[V];
};

template <typename T>
__device__ __forceinline__ T lane_prev(T x, T halo, int lane) {
    T y = __shfl_up(x, 1, 64);
    return lane == 0 ? halo : y;
}
template <typename T>
__device__ __forceinline__ T lane_next(T x, T halo, int lane) {
    T y = __shfl_down(x, 1, 64);
    return lane == 63 ? halo : y;
}

// Neumaier-compensated accumulation.
__device__ __forceinline__ void kahan_add(double& s, double& c, double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

__device__ __forceinline__ unsigned long long key_neg_max() { return fkey(-DBL_MAX); }

// One row of cells (and halos) of one array at row r.
template <typename T, int V>
__device__ __forceinline__ void load_row(T (&dst)[V], const T* __restrict__ base, long long P,
                                         int r, int rows, int c0) {
    if (r >= 0 && r < rows && c0 < P) {
        Pack<T, V> p = *reinterpret_cast<const Pack<T, V>*>(base + (long long)r * P + c0);
#pragma unroll
        for (int v = 0; v < V; v++) dst[v] = p.v[v];
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) dst[v] = T(0);
    }
}
template <typename T>
__device__ __forceinline__ T load_one(const T* __restrict__ base, long long P, int r, int rows,
                                      int col, bool ok) {
    return (ok && r >= 0 && r < rows) ? base[(long long)r * P + col] : T(0);
}

template <typename T, int V, bool F64>
struct Stepper {
    static constexpr int W = 64 * V;

    // Interior face flux; q values are already scaled to q_old units.
    static __device__ __forceinline__ T flux(T z0, T z1, T h0, T h1, T q0, T qt, T nod,
                                             const FluxConstsF64& kd, const FluxConstsF32& kf) {
        if (z0 == nod || z1 == nod) return T(0);
        if constexpr (F64) {
            return face_flux_ref_t1(z0, z1, h0, h1, q0, qt, kd);
        } else {
            return face_flux_f32(z0, z1, h0, h1, q0, qt, kf);
        }
    }

    static __device__ __forceinline__ void run(const ScnDev& S, const TileDesc td, int kmod6, int lane) {
        const int cur = kmod6 & 1, nxt = cur ^ 1;
        const int hs_cur = kmod6 % 3, hs_nxt = (kmod6 + 1) % 3, hs_rst = (kmod6 + 2) % 3;
        Ctl* ctl = S.ctl;
        const int tile_local = td.strip + td.seg * S.nstrips;

        const double t = ctl->t[cur];
        const double next_mass = ctl->next_mass[cur];
        const double hmax_in = fkey_inv_d(ctl->hmax[hs_cur]);
        const bool skip = !(t < S.sim_time);  // main.cpp:227
        const double dt = cfl_dt(hmax_in, S.thresh, S.cfl, S.dx, S.g, S.init_tstep);
        const bool report_due = !skip && (t + dt >= next_mass);

        if (td.first) bookkeeping(S, cur, nxt, hs_cur, hs_nxt, hs_rst, t, dt, next_mass, skip,
                                  report_due, lane);
        if (skip) return;

        const int nx = S.nx, ny = S.ny;
        const long long P = S.P;
        const int i0 = td.strip * W;
        const int j0 = td.seg * S.tile_rows;
        const int j1 = min(j0 + S.tile_rows, ny);
        const int c0 = i0 + lane * V;
        const T nod = (T)S.nodata_t;
        const T* __restrict__ z = reinterpret_cast<const T*>(S.z);
        const T* __restrict__ Ho = reinterpret_cast<const T*>(S.H[cur]);
        T* __restrict__ Hn = reinterpret_cast<T*>(S.H[nxt]);
        const T* __restrict__ Qxo = reinterpret_cast<const T*>(S.Qx[cur]);
        T* __restrict__ Qxn = reinterpret_cast<T*>(S.Qx[nxt]);
        const T* __restrict__ Qyo = reinterpret_cast<const T*>(S.Qy[cur]);
        T* __restrict__ Qyn = reinterpret_cast<T*>(S.Qy[nxt]);
        const T sc = (T)S.dxinv;  // Q -> q_old (kernels.cpp:8-16)
        const T dxT = (T)S.dx;
        const T gT = (T)S.g;

        FluxConstsF64 kd{S.dx, S.fn, S.thresh, S.maxh, S.g * dt, S.theta};
        FluxConstsF32 kf{(float)S.dx, (float)S.thresh, (float)S.maxh,
                         (float)(S.g * dt / S.dx), (float)(S.g * dt * S.fn * S.fn)};
        const double dA = S.dA;
        const T dtdA = (T)(dt / dA);

        // rain depth of this step (main.cpp:265-268, kernels.cpp:356-358)
        double rain_depth = 0.0;
        bool rain_on = false;
        if (S.rain_enabled) {
            double r = interp_series(S.rain_t, S.rain_r, S.n_rain, 0.0, t);
            rain_on = (r > 0.0);
            if (rain_on) rain_depth = r * dt;
        }
        const T rainT = (T)rain_depth;

        // halo columns: lane 0 reads column i0-1, every other lane column i0+W
        // (only lane 63 uses it; the others hit the same cache line)
        const int hcol = (lane == 0) ? i0 - 1 : i0 + W;
        const bool hcell = (lane == 0) ? (i0 > 0) : (i0 + W < nx);   // a cell column
        const bool hface = (lane == 0) ? false : (i0 + W <= nx);     // an x-face column
        const int hcol_c = hcell ? hcol : 0;
        const int hcol_f = hface ? hcol : 0;

        // sliding window: c = row j, s = row j+1, p = row j+2 (prefetch)
        T hc[V], zc[V], qxc[V], qyc[V];
        T hs[V], zs[V], qxs[V], qys[V];
        T hH_c, zH_c, qyH_c, qxH_c;  // halo values for row j
        T hH_s, zH_s, qyH_s, qxH_s;  // halo values for row j+1

        load_row<T, V>(hc, Ho, P, j0, ny, c0);
        load_row<T, V>(zc, z, P, j0, ny, c0);
        load_row<T, V>(qxc, Qxo, P, j0, ny, c0);
        load_row<T, V>(qyc, Qyo, P, j0, ny + 1, c0);
        hH_c = load_one(Ho, P, j0, ny, hcol_c, hcell);
        zH_c = load_one(z, P, j0, ny, hcol_c, hcell);
        qyH_c = load_one(Qyo, P, j0, ny + 1, hcol_c, hcell);
        qxH_c = load_one(Qxo, P, j0, ny, hcol_f, hface);
        load_row<T, V>(hs, Ho, P, j0 + 1, ny, c0);
        load_row<T, V>(zs, z, P, j0 + 1, ny, c0);
        load_row<T, V>(qxs, Qxo, P, j0 + 1, ny, c0);
        load_row<T, V>(qys, Qyo, P, j0 + 1, ny + 1, c0);
        hH_s = load_one(Ho, P, j0 + 1, ny, hcol_c, hcell);
        zH_s = load_one(z, P, j0 + 1, ny, hcol_c, hcell);
        qyH_s = load_one(Qyo, P, j0 + 1, ny + 1, hcol_c, hcell);
        qxH_s = load_one(Qxo, P, j0 + 1, ny, hcol_f, hface);

        // north faces of row j0
        T qyN[V];
        if (j0 == 0) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                T q = T(0);
                if (S.north_free && c0 + v < nx) q = edge_out_neg<T>(hc[v], zc[v] == nod, dxT, gT);
                qyN[v] = q;
            }
            if (c0 < P) {
                Pack<T, V> o;
#pragma unroll
                for (int v = 0; v < V; v++) o.v[v] = (c0 + v < nx) ? qyN[v] : T(0);
                *reinterpret_cast<Pack<T, V>*>(Qyn + c0) = o;
            }
        } else {
            T ha[V], za[V], qxa[V];
            load_row<T, V>(ha, Ho, P, j0 - 1, ny, c0);
            load_row<T, V>(za, z, P, j0 - 1, ny, c0);
            load_row<T, V>(qxa, Qxo, P, j0 - 1, ny, c0);
            T qxH_a = load_one(Qxo, P, j0 - 1, ny, hcol_f, hface);
            T qxa_n = lane_next(qxa[0], qxH_a, lane);
            T qxc_n = lane_next(qxc[0], qxH_c, lane);
#pragma unroll
            for (int v = 0; v < V; v++) {
                T a1 = (v + 1 < V) ? qxa[v + 1] : qxa_n;
                T c1 = (v + 1 < V) ? qxc[v + 1] : qxc_n;
                T qt = (((qxa[v] * sc + a1 * sc) + qxc[v] * sc) + c1 * sc) * T(0.25);
                qyN[v] = (c0 + v < nx) ? flux(za[v], zc[v], ha[v], hc[v], qyc[v] * sc, qt, nod, kd, kf)
                                       : T(0);
            }
        }

        // per-tile point lists (uniform)
        int sp = S.src_begin[tile_local];
        const int se = S.src_begin[tile_local + 1];
        int src_next_j = (sp < se) ? S.src_j[sp] : INT_MAX;
        int op = S.ovr_begin[tile_local];
        const int oe = S.ovr_begin[tile_local + 1];
        int ovr_next_j = (op < oe) ? S.ovr_j[op] : INT_MAX;

        T hmax_l = F64 ? T(-DBL_MAX) : T(-FLT_MAX);
        double vol_s = 0.0, vol_c = 0.0;
        unsigned long long wet = 0;

#pragma unroll 1
        for (int j = j0; j < j1; j++) {
            // prefetch row j+2
            T hp[V], zp[V], qxp[V], qyp[V];
            load_row<T, V>(hp, Ho, P, j + 2, ny, c0);
            load_row<T, V>(zp, z, P, j + 2, ny, c0);
            load_row<T, V>(qxp, Qxo, P, j + 2, ny, c0);
            load_row<T, V>(qyp, Qyo, P, j + 2, ny + 1, c0);
            T hH_p = load_one(Ho, P, j + 2, ny, hcol_c, hcell);
            T zH_p = load_one(z, P, j + 2, ny, hcol_c, hcell);
            T qyH_p = load_one(Qyo, P, j + 2, ny + 1, hcol_c, hcell);
            T qxH_p = load_one(Qxo, P, j + 2, ny, hcol_f, hface);

            // neighbours across lanes for rows j and j+1
            const T h_prev = lane_prev(hc[V - 1], hH_c, lane);
            const T z_prev = lane_prev(zc[V - 1], zH_c, lane);
            const T qyc_prev = lane_prev(qyc[V - 1], qyH_c, lane);
            const T qys_prev = lane_prev(qys[V - 1], qyH_s, lane);
            const T qxc_next = lane_next(qxc[0], qxH_c, lane);
            const T qxs_next = lane_next(qxs[0], qxH_s, lane);

            // x faces f = c0+v (west face of column c0+v)
            T qxw[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                const int f = c0 + v;
                T zl = (v == 0) ? z_prev : zc[v - 1];
                T hl = (v == 0) ? h_prev : hc[v - 1];
                T ql0 = (v == 0) ? qyc_prev : qyc[v - 1];
                T ql1 = (v == 0) ? qys_prev : qys[v - 1];
                T q;
                if (f > 0 && f < nx) {
                    T qt = (((ql0 * sc + qyc[v] * sc) + ql1 * sc) + qys[v] * sc) * T(0.25);
                    q = flux(zl, zc[v], hl, hc[v], qxc[v] * sc, qt, nod, kd, kf);
                } else if (f == 0) {
                    q = S.west_free ? edge_west<T>(zc[0], hc[0], zc[1], hc[1], zc[0] == nod, dxT, gT)
                                    : T(0);
                } else if (f == nx) {
                    q = S.east_free ? edge_out_pos<T>(hl, zl == nod, dxT, gT) : T(0);
                } else {
                    q = T(0);
                }
                qxw[v] = q;
            }
            // east face of the lane's last column: the next lane's west face,
            // or for lane 63 the strip's east face
            T qxe_last = __shfl_down(qxw[0], 1, 64);
            if (lane == 63) {
                const int f = i0 + W;
                T q = T(0);
                if (f < nx) {
                    T qt = (((qyc[V - 1] * sc + qyH_c * sc) + qys[V - 1] * sc) + qyH_s * sc) * T(0.25);
                    q = flux(zc[V - 1], zH_c, hc[V - 1], hH_c, qxH_c * sc, qt, nod, kd, kf);
                } else if (f == nx) {
                    q = S.east_free ? edge_out_pos<T>(hc[V - 1], zc[V - 1] == nod, dxT, gT) : T(0);
                }
                qxe_last = q;
            }

            // y faces south of row j (face row j+1)
            T qyS[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                T q = T(0);
                if (c0 + v < nx) {
                    if (j + 1 < ny) {
                        T c1 = (v + 1 < V) ? qxc[v + 1] : qxc_next;
                        T s1 = (v + 1 < V) ? qxs[v + 1] : qxs_next;
                        T qt = (((qxc[v] * sc + c1 * sc) + qxs[v] * sc) + s1 * sc) * T(0.25);
                        q = flux(zc[v], zs[v], hc[v], hs[v], qys[v] * sc, qt, nod, kd, kf);
                    } else if (S.south_free) {
                        q = edge_out_pos<T>(hc[v], zc[v] == nod, dxT, gT);
                    }
                }
                qyS[v] = q;
            }

            // cell update of row j
            T hn[V];
#pragma unroll
            for (int v = 0; v < V; v++) hn[v] = hc[v];
            // point sources of this row, in (column, file) order (main.cpp:245-261)
            while (src_next_j == j) {
                const int col = S.src_i[sp];
                const int pt = S.src_pt[sp];
                double val;
                if (S.pt_type[pt] == 4) val = S.pt_value[pt];
                else val = interp_series(S.series_t + S.pt_off[pt], S.series_v + S.pt_off[pt],
                                         S.pt_len[pt], S.pt_value[pt], t);
                const double ps = val * S.dx * dt / dA;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    if (col == c0 + v) {
                        T h = hn[v] + (T)ps;
                        if (h < T(0)) h = T(0);
                        hn[v] = h;
                    }
                }
                sp++;
                src_next_j = (sp < se) ? S.src_j[sp] : INT_MAX;
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                T h = hn[v];
                if (rain_on && zc[v] != nod) {  // kernels.cpp:361-364
                    h = h + rainT;
                    if (h < T(0)) h = T(0);
                }
                const T qe = (v + 1 < V) ? qxw[v + 1] : qxe_last;
                const T div = ((qxw[v] - qe) + qyN[v]) - qyS[v];
                if constexpr (F64) h = h + (T)(dt * (double)div / dA);  // kernels.cpp:286
                else h = h + dtdA * div;
                if (h < T(0)) h = T(0);
                hn[v] = h;
            }
            // FREE / HFIX / HVAR overrides, last in file order wins
            while (ovr_next_j == j) {
                const int col = S.ovr_i[op];
                const int pt = S.ovr_pt[op];
                const int ty = S.pt_type[pt];
                double val = 0.0;
                if (ty == 2) val = S.pt_value[pt];
                else if (ty == 3)
                    val = interp_series(S.series_t + S.pt_off[pt], S.series_v + S.pt_off[pt],
                                        S.pt_len[pt], S.pt_value[pt], t);
#pragma unroll
                for (int v = 0; v < V; v++) {
                    if (col == c0 + v) hn[v] = (T)val;
                }
                op++;
                ovr_next_j = (op < oe) ? S.ovr_j[op] : INT_MAX;
            }
            // H maximum of the new state over valid cells, mass-report sums
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (c0 + v < nx) {
                    if (zc[v] != nod && hn[v] > hmax_l) hmax_l = hn[v];
                    if (report_due) {
                        const double hd = (double)hn[v];
                        kahan_add(vol_s, vol_c, hd * dA);
                        if (hd > 0.01) wet++;
                    }
                }
            }

            // stores: H row j, Qx faces of row j, Qy face row j+1
            if (c0 < P) {
                Pack<T, V> oh, ox, oy;
#pragma unroll
                for (int v = 0; v < V; v++) {
                    oh.v[v] = (c0 + v < nx) ? hn[v] : T(0);
                    ox.v[v] = (c0 + v <= nx) ? qxw[v] : T(0);
                    oy.v[v] = (c0 + v < nx) ? qyS[v] : T(0);
                }
                const long long row = (long long)j * P + c0;
                *reinterpret_cast<Pack<T, V>*>(Hn + row) = oh;
                *reinterpret_cast<Pack<T, V>*>(Qxn + row) = ox;
                *reinterpret_cast<Pack<T, V>*>(Qyn + row + P) = oy;
            }
            if (lane == 63 && i0 + W == nx) Qxn[(long long)j * P + nx] = qxe_last;

            // slide the window
#pragma unroll
            for (int v = 0; v < V; v++) {
                qyN[v] = qyS[v];
                hc[v] = hs[v]; zc[v] = zs[v]; qxc[v] = qxs[v]; qyc[v] = qys[v];
                hs[v] = hp[v]; zs[v] = zp[v]; qxs[v] = qxp[v]; qys[v] = qyp[v];
            }
            hH_c = hH_s; zH_c = zH_s; qyH_c = qyH_s; qxH_c = qxH_s;
            hH_s = hH_p; zH_s = zH_p; qyH_s = qyH_p; qxH_s = qxH_p;
        }

        // wavefront max, one atomic per wavefront
        for (int off = 32; off > 0; off >>= 1) hmax_l = getmax(hmax_l, __shfl_xor(hmax_l, off, 64));
        if (lane == 0) {
            // the empty-set identity stays the lowest double, as a Kokkos max reducer's
            const double m = (hmax_l == (F64 ? T(-DBL_MAX) : T(-FLT_MAX))) ? -DBL_MAX : (double)hmax_l;
            atomicMax(&ctl->hmax[hs_nxt], fkey(m));
        }
        if (report_due) {
            for (int off = 32; off > 0; off >>= 1) {
                double os = __shfl_xor(vol_s, off, 64);
                double oc = __shfl_xor(vol_c, off, 64);
                unsigned long long ow = __shfl_xor(wet, off, 64);
                // lanes pair up symmetrically; both keep the same merged pair
                double a = (lane & off) ? os : vol_s;
                double b = (lane & off) ? vol_s : os;
                double ca = (lane & off) ? oc : vol_c;
                double cb = (lane & off) ? vol_c : oc;
                double s = a, c = ca + cb;
                kahan_add(s, c, b);
                vol_s = s;
                vol_c = c;
                wet += ow;
            }
            if (lane == 0) {
                const long long slot = (long long)cur * S.ntiles + tile_local;
                S.part_vol[slot] = vol_s;
                S.part_comp[slot] = vol_c;
                S.part_area[slot] = wet;
            }
        }
    }

    // Scenario bookkeeping, done by the wavefront of the scenario's first tile:
    // finalise the previous step's mass report, advance t, rotate the H-max
    // slots, log dt.
    static __device__ __forceinline__ void bookkeeping(const ScnDev& S, int cur, int nxt, int hs_cur, int hs_nxt,
                                       int hs_rst, double t, double dt, double next_mass,
                                       bool skip, bool report_due, int lane) {
        Ctl* ctl = S.ctl;
        finalize_report(S, lane);
        if (lane == 0) {
            if (!skip) {
                const double t_new = t + dt;  // main.cpp:287
                ctl->t[nxt] = t_new;
                ctl->hmax[hs_rst] = key_neg_max();
                const long long n = ctl->n_done;
                if (n < S.dt_log_cap) S.dt_log[n] = dt;
                ctl->n_done = n + 1;
                ctl->last_dt = dt;
                if (report_due) {  // main.cpp:291-304
                    ctl->next_mass[nxt] = next_mass + S.mass_int;
                    ctl->pending = 1;
                    ctl->pending_buf = cur;
                    ctl->pending_t = t_new;
                    ctl->pending_dt = dt;
                } else {
                    ctl->next_mass[nxt] = next_mass;
                }
            } else {
                ctl->t[nxt] = t;
                ctl->next_mass[nxt] = next_mass;
                ctl->hmax[hs_nxt] = ctl->hmax[hs_cur];
            }
        }
    }

    static __device__ __forceinline__ void finalize_report(const ScnDev& S, int lane) {
        Ctl* ctl = S.ctl;
        if (!ctl->pending) return;
        const long long base = (long long)ctl->pending_buf * S.ntiles;
        double s = 0.0, c = 0.0;
        unsigned long long wet = 0;
        for (int i = lane; i < S.ntiles; i += 64) {
            kahan_add(s, c, S.part_vol[base + i]);
            c += S.part_comp[base + i];
            wet += S.part_area[base + i];
        }
        for (int off = 32; off > 0; off >>= 1) {
            double os = __shfl_xor(s, off, 64);
            double oc = __shfl_xor(c, off, 64);
            unsigned long long ow = __shfl_xor(wet, off, 64);
            double a = (lane & off) ? os : s;
            double b = (lane & off) ? s : os;
            double ca = (lane & off) ? oc : c;
            double cb = (lane & off) ? c : oc;
            double ss = a, cc = ca + cb;
            kahan_add(ss, cc, b);
            s = ss;
            c = cc;
            wet += ow;
        }
        if (lane == 0) {
            const int r = ctl->n_reports;
            if (r < S.report_cap) {
                double* rec = S.reports + 4LL * r;
                rec[0] = ctl->pending_t;
                rec[1] = ctl->pending_dt;
                rec[2] = s + c;
                rec[3] = (double)wet * S.dA;
            }
            ctl->n_reports = r + 1;
            ctl->pending = 0;
        }
    }
};

// Maps hardware block ids to logical ones so that each XCD (blocks are
// dealt round-robin to the 8 XCDs) receives one contiguous range of tiles:
// neighbouring strips then share an L2.
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

template <typename T, int V, bool F64, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_step(const ScnDev* __restrict__ scns, const TileDesc* __restrict__ tiles, int ntiles, int kmod6) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tile = lb * WPB + wv;
    if (tile >= ntiles) return;
    const TileDesc td = tiles[tile];
    Stepper<T, V, F64>::run(scns[td.scn], td, kmod6, lane);
}

}  // namespace af
