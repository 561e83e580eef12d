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
    int par[2];                   // buffer holding the state entering step k: par[k & 1]
    // AF_SCN_* bits; non-zero stops the scenario for good.  Step k reads
    // status[k & 1] and sets bits in status[(k + 1) & 1] (the first tile's
    // wave carries the old bits over), so no wave of a step sees a bit that
    // another wave of the same step set: every wave of a scenario takes the
    // same skip decision.
    int status[2];
    // cells deeper than 0.01 m at the last finalized mass report (the
    // initial state's before the first), for the host's choice of variant
    unsigned long long last_wet;
};

// Ctl.status bits (the values of AF_SCN_NONFINITE / AF_SCN_STALLED).
constexpr int kStatusNonfinite = 1;   // a step produced a NaN or infinite depth
constexpr int kStatusStalled = 2;     // t + dt == t: the clock cannot advance

// Per-scenario constants and device pointers.
struct ScnDev {
    int nx, ny;
    long long P;                  // row pitch in elements (all state arrays)
    int nstrips, tile_rows;
    int tile_base;                // first global tile of this scenario
    int ntiles;
    void* z;                      // Zt<T>[ny*P]: FP64 elevations (NODATA cells hold
                                  // nodata_t), or FP32-path fixed point (INT_MIN)
    void* H[2];                   // T[ny*P]
    void* Qx[2];                  // T[ny*P], face f of row j at j*P + f, f in [0, nx]
    void* Qy[2];                  // T[(ny+1)*P], face row g (north of row g) at g*P + i
    double dx, dxinv, dA, fn, thresh, maxh, init_tstep, cfl, g, theta, sim_time, mass_int;
    double nodata_t;              // the FP64 terrain's NODATA value
    double zscale;                // metres per fixed-point unit of the FP32 terrain
    double zbase;                 // elevation of fixed-point zero
    int west_free, east_free, north_free, south_free;
    // point sources (QFIX/QVAR) and overrides (FREE/HFIX/HVAR), sorted per
    // tile by (row, column, file order); *_begin has ntiles+1 entries.  An
    // entry's value is its series [off, off+len) at the step's start time,
    // or `val` when the series is empty (constants have len = 0).
    const int* src_begin;
    const int* src_j;
    const int* src_i;
    const double* src_val;
    const int* src_off;
    const int* src_len;
    const int* ovr_begin;
    const int* ovr_j;
    const int* ovr_i;
    const double* ovr_val;
    const int* ovr_off;
    const int* ovr_len;
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

// V adjacent values of one row, moved as one 8- or 16-byte access
template <typename T, int V>
using Vec = T __attribute__((ext_vector_type(V)));

// Pointers loaded from the scenario table are generic; the kernel works on
// them as global-memory pointers so that loads and stores are global_* (not
// flat_*) instructions and uniform loads can go through the scalar unit.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
using lptr = __attribute__((address_space(3))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glob(const void* p) {
    return (gptr<T>)(const_cast<void*>(p));
}

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

// One row of one array at row r (r >= 0), V columns from c0.  The load is
// always issued, from a clamped in-bounds address: a load under a branch
// would leave the wait counter unknown and make the next use wait for every
// load in flight.  Values read for rows or columns outside the array are
// never consumed: every face and cell that would use them is masked by its
// own row or column test.
template <typename T, int V>
__device__ __forceinline__ void load_row(T (&dst)[V], gptr<const T> base, long long P,
                                         int r, int rows, int c0) {
    const int rr = (r < rows) ? r : rows - 1;
    const int cc = (c0 < P) ? c0 : 0;
    Vec<T, V> p = *(gptr<const Vec<T, V>>)(base + (long long)rr * P + cc);
#pragma unroll
    for (int v = 0; v < V; v++) dst[v] = p[v];
}
// One value at (r, col); col must be a valid column.
template <typename T>
__device__ __forceinline__ T load_one(gptr<const T> base, long long P, int r, int rows, int col) {
    const int rr = (r < rows) ? r : rows - 1;
    return base[(long long)rr * P + col];
}

// Terrain element of a state precision: FP64 keeps the elevations as given;
// FP32 keeps them as 32-bit fixed point above a base (integer units of a
// power-of-two step, NODATA = INT_MIN), so that every face's terrain step is
// an exact integer difference (terrain_step).
template <typename T>
struct TerrainOf {
    using type = T;
};
template <>
struct TerrainOf<float> {
    using type = int;
};
template <typename T>
using Zt = typename TerrainOf<T>::type;

// One row of the sliding window: V columns of H, z, Qx and Qy per lane, plus
// the strip-edge value of each (lane 0: column i0-1; lane 63: column i0+W).
template <typename T, int V>
struct RowData {
    T h[V];
    Zt<T> z[V];
    T qx[V], qy[V];
    T hH;
    Zt<T> zH;
    T qyH, qxH;
};

// Results of one row before they are stored: H of the row, the west x-face
// of each column, the south y-face of each column, and lane 63's strip-edge
// east face.
template <typename T, int V>
struct RowOut {
    T h[V], qx[V], qy[V];
    T qxe;
};

// Per-tile constants of one step.
template <typename T>
struct TileCtx {
    gptr<const Zt<T>> z;
    gptr<const T> Ho, Qxo, Qyo;
    gptr<T> Hn, Qxn, Qyn;
    long long P;
    int nx, ny, i0, j0, j1, c0, rc, rf, lane;
    int hcol_c, hcol_f;
    bool hcell, hface;
    Zt<T> nod;       // the terrain's NODATA value
    T znod;          // FP32: the NODATA value in metres above the terrain base
    T sc, dxT, gT, dtdA, rainT;
    bool rain_on, report_due;
    double t, dt, dA;
    FluxConstsF64 kd;
    FluxConstsF32 kf;
    lptr<int> lrc;   // this wave's point batches in LDS: sources first, then overrides;
    lptr<T> lval;    // packed row/column and the value applied
};

// Per-tile values carried from row to row.
//
// Point lists.  At the tile's start, the first kPointBatch sources (and,
// separately, overrides) of the tile are evaluated across the lanes, one
// entry per lane and round, and their packed row and column,
// (row - j0) << 9 | (column - i0), and the value they apply go to the wave's
// LDS slice.  The row loop applies them from LDS in list order, and takes any
// entries past the batch one at a time with uniform loads.  The cursors are
// uniform and live in scalar registers; the row of the next batched entry
// comes from LDS, so no row waits on the vector loads in flight to decide
// whether a point falls on it.
constexpr int kPointBatch = 128;

template <typename T, int V>
struct TileState {
    T qyN[V];
    int sp, se, sk, sn, src_next_j;   // sources: first, end, next (from sp), batched; next row
    int op, oe, ok, on, ovr_next_j;   // overrides: likewise
    T hmax;
    bool finite;                  // every depth this lane produced is finite
    double vol_s, vol_c;
    unsigned long long wet;
};

template <typename T, int V, bool F64, int PF = 1>
struct Stepper {
    static constexpr int W = 64 * V;

    // Interior face flux from stored discharges: Q0 of the face and the four
    // Q of its transverse average (kernels.cpp:79, 129).
    static __device__ __forceinline__ T flux(Zt<T> z0, Zt<T> z1, T h0, T h1, T Q0, T a, T b, T c,
                                             T d, Zt<T> nod, T sc, const FluxConstsF64& kd,
                                             const FluxConstsF32& kf) {
        if constexpr (F64) {
            if (z0 == nod || z1 == nod) return T(0);
            return face_flux_ref_t1_q(z0, z1, h0, h1, Q0, a, b, c, d, sc, kd);
        } else {
            // no branch (see face_flux_f32)
            const T q = face_flux_f32_q(terrain_step(z0, z1, kf.zsc), h0, h1, Q0, a, b, c, d, kf);
            return (z0 == nod || z1 == nod) ? T(0) : q;
        }
    }

    // West edge face of row c (kernels.cpp:155-179)
    static __device__ __forceinline__ T west_edge(const TileCtx<T>& x, const RowData<T, V>& c) {
        if constexpr (F64)
            return edge_west<T>(c.z[0], c.h[0], c.z[1], c.h[1], c.z[0] == x.nod, x.dxT, x.gT);
        else {
            // the inner neighbour enters without a NODATA test (kernels.cpp:162-164):
            // a NODATA neighbour stands at the NODATA value itself
            const float dz10 = (c.z[1] == x.nod)
                                   ? x.znod - (float)c.z[0] * x.kf.zsc
                                   : terrain_step(c.z[1], c.z[0], x.kf.zsc);
            return edge_west_f32(c.h[0], c.h[1], dz10, c.z[0] == x.nod, x.dxT, x.gT);
        }
    }

    // Row r of H, z, Qx (cell rows, r < rc) and Qy (face rows, r < rf).
    static __device__ __forceinline__ void load(RowData<T, V>& d, const TileCtx<T>& x, int r) {
        load_row<T, V>(d.h, x.Ho, x.P, r, x.rc, x.c0);
        load_row<Zt<T>, V>(d.z, x.z, x.P, r, x.rc, x.c0);
        load_row<T, V>(d.qx, x.Qxo, x.P, r, x.rc, x.c0);
        load_row<T, V>(d.qy, x.Qyo, x.P, r, x.rf, x.c0);
        d.hH = load_one(x.Ho, x.P, r, x.rc, x.hcol_c);
        d.zH = load_one(x.z, x.P, r, x.rc, x.hcol_c);
        d.qyH = load_one(x.Qyo, x.P, r, x.rf, x.hcol_c);
        d.qxH = load_one(x.Qxo, x.P, r, x.rc, x.hcol_f);
    }

    // Row j of the tile from c (row j) and s (row j+1): faces, sources, rain,
    // depth update, overrides and the H-max / report folds; results in o.
    template <bool EDGE>
    static __device__ __forceinline__ void compute(const ScnDev& S, const TileCtx<T>& x,
                                                   TileState<T, V>& st, const RowData<T, V>& c,
                                                   const RowData<T, V>& s, int j,
                                                   RowOut<T, V>& o) {
        const int lane = x.lane, c0 = x.c0, nx = x.nx;
        const Zt<T> nod = x.nod;
        const T sc = x.sc;

        // neighbours across lanes for rows j and j+1
        const T h_prev = lane_prev(c.h[V - 1], c.hH, lane);
        const Zt<T> z_prev = lane_prev(c.z[V - 1], c.zH, lane);
        const T qyc_prev = lane_prev(c.qy[V - 1], c.qyH, lane);
        const T qys_prev = lane_prev(s.qy[V - 1], s.qyH, lane);
        const T qxc_next = lane_next(c.qx[0], c.qxH, lane);
        const T qxs_next = lane_next(s.qx[0], s.qxH, lane);

        // x faces f = c0+v (west face of column c0+v)
        T qxw[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int f = c0 + v;
            Zt<T> zl = (v == 0) ? z_prev : c.z[v - 1];
            T hl = (v == 0) ? h_prev : c.h[v - 1];
            T ql0 = (v == 0) ? qyc_prev : c.qy[v - 1];
            T ql1 = (v == 0) ? qys_prev : s.qy[v - 1];
            T q;
            if (!EDGE || (f > 0 && f < nx)) {
                q = flux(zl, c.z[v], hl, c.h[v], c.qx[v], ql0, c.qy[v], ql1, s.qy[v], nod, sc, x.kd,
                         x.kf);
            } else if (f == 0) {
                q = S.west_free ? west_edge(x, c) : T(0);
            } else if (f == nx) {
                q = S.east_free ? edge_out_pos<T>(hl, zl == nod, x.dxT, x.gT) : T(0);
            } else {
                q = T(0);
            }
            qxw[v] = q;
        }
        // east face of the lane's last column: the next lane's west face,
        // or for lane 63 the strip's east face.  Every lane evaluates the
        // strip's east face (a face evaluated for one lane occupies the
        // wavefront as long as one for all) and lane 63 keeps it.
        T qxe_last = __shfl_down(qxw[0], 1, 64);
        {
            const int f = x.i0 + W;
            T q = T(0);
            if (!EDGE || f < nx) {
                q = flux(c.z[V - 1], c.zH, c.h[V - 1], c.hH, c.qxH, c.qy[V - 1], c.qyH, s.qy[V - 1],
                         s.qyH, nod, sc, x.kd, x.kf);
            } else if (f == nx) {
                q = S.east_free ? edge_out_pos<T>(c.h[V - 1], c.z[V - 1] == nod, x.dxT, x.gT)
                                : T(0);
            }
            if (lane == 63) qxe_last = q;
        }

        // y faces south of row j (face row j+1)
        T qyS[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            T q = T(0);
            if (!EDGE || c0 + v < nx) {
                if (j + 1 < x.ny) {
                    T c1 = (v + 1 < V) ? c.qx[v + 1] : qxc_next;
                    T s1 = (v + 1 < V) ? s.qx[v + 1] : qxs_next;
                    q = flux(c.z[v], s.z[v], c.h[v], s.h[v], s.qy[v], c.qx[v], c1, s.qx[v], s1, nod,
                             sc, x.kd, x.kf);
                } else if (S.south_free) {
                    q = edge_out_pos<T>(c.h[v], c.z[v] == nod, x.dxT, x.gT);
                }
            }
            qyS[v] = q;
        }

        // cell update of row j
        T hn[V];
#pragma unroll
        for (int v = 0; v < V; v++) hn[v] = c.h[v];
        // point sources of this row, in (column, file) order (main.cpp:245-261)
        while (__builtin_expect(st.src_next_j == j, 0)) apply_source(S, x, st, hn);
#pragma unroll
        for (int v = 0; v < V; v++) {
            T h = hn[v];
            if (x.rain_on && c.z[v] != nod) {  // kernels.cpp:361-364
                h = h + x.rainT;
                if (h < T(0)) h = T(0);
            }
            const T qe = (v + 1 < V) ? qxw[v + 1] : qxe_last;
            const T div = ((qxw[v] - qe) + st.qyN[v]) - qyS[v];
            if constexpr (F64) h = h + (T)(x.dt * (double)div / x.dA);  // kernels.cpp:286
            else h = h + x.dtdA * div;
            if (h < T(0)) h = T(0);
            hn[v] = h;
        }
        // FREE / HFIX / HVAR overrides, last in file order wins
        while (__builtin_expect(st.ovr_next_j == j, 0)) apply_override(S, x, st, hn);
        // H maximum of the new state over valid cells, mass-report sums
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (!EDGE || c0 + v < nx) {
                if (c.z[v] != nod && hn[v] > st.hmax) st.hmax = hn[v];
                // NaN never wins the maximum (as in kernels.cpp:32), so it is
                // looked for separately
                st.finite = st.finite && __builtin_isfinite(hn[v]);
                if (x.report_due) {
                    const double hd = (double)hn[v];
                    kahan_add(st.vol_s, st.vol_c, hd * x.dA);
                    if (hd > 0.01) st.wet++;
                }
            }
        }

#pragma unroll
        for (int v = 0; v < V; v++) {
            o.h[v] = hn[v];
            o.qx[v] = qxw[v];
            o.qy[v] = qyS[v];
            st.qyN[v] = qyS[v];
        }
        o.qxe = qxe_last;
    }

    // Stores row j: H of the row, its x faces and the y faces south of it.
    template <bool EDGE>
    static __device__ __forceinline__ void store(const TileCtx<T>& x, const RowOut<T, V>& o,
                                                 int j) {
        const int c0 = x.c0, nx = x.nx;
        if (!EDGE || c0 < x.P) {
            Vec<T, V> oh, ox, oy;
#pragma unroll
            for (int v = 0; v < V; v++) {
                oh[v] = (!EDGE || c0 + v < nx) ? o.h[v] : T(0);
                ox[v] = (!EDGE || c0 + v <= nx) ? o.qx[v] : T(0);
                oy[v] = (!EDGE || c0 + v < nx) ? o.qy[v] : T(0);
            }
            // streaming stores: the new state is not read again in this
            // launch, so it is written with the non-temporal hint (measured
            // +4% at FP32, +2% at FP64 on the 16384 x 16384 step)
            const long long rowo = (long long)j * x.P + c0;
            __builtin_nontemporal_store(oh, (gptr<Vec<T, V>>)(x.Hn + rowo));
            __builtin_nontemporal_store(ox, (gptr<Vec<T, V>>)(x.Qxn + rowo));
            __builtin_nontemporal_store(oy, (gptr<Vec<T, V>>)(x.Qyn + rowo + x.P));
        }
        if (EDGE && x.lane == 63 && x.i0 + W == nx) x.Qxn[(long long)j * x.P + nx] = o.qxe;
    }

    // One row per call: row j+2 is loaded into nw first, so that one row of
    // loads is in flight while row j computes.
    template <bool EDGE>
    static __device__ __forceinline__ void row(const ScnDev& S, const TileCtx<T>& x,
                                               TileState<T, V>& st, const RowData<T, V>& c,
                                               const RowData<T, V>& s, RowData<T, V>& nw, int j) {
        load(nw, x, j + 2);
        RowOut<T, V> o;
        compute<EDGE>(S, x, st, c, s, j, o);
        store<EDGE>(x, o, j);
    }

    // Two rows per call: rows j+2 and j+3 are loaded as one burst, rows j and
    // j+1 computed, and their six stores issued together.
    template <bool EDGE>
    static __device__ __forceinline__ void row2(const ScnDev& S, const TileCtx<T>& x,
                                                TileState<T, V>& st, const RowData<T, V>& a,
                                                const RowData<T, V>& b, RowData<T, V>& c,
                                                RowData<T, V>& d, int j) {
        load(c, x, j + 2);
        load(d, x, j + 3);
        RowOut<T, V> o0, o1;
        compute<EDGE>(S, x, st, a, b, j, o0);
        const bool two = (j + 1 < x.j1);
        if (two) compute<EDGE>(S, x, st, b, c, j + 1, o1);
        store<EDGE>(x, o0, j);
        if (two) store<EDGE>(x, o1, j + 1);
    }

    // Value of source entry k (the list is sorted per tile): v at the step's
    // start time, then ps = v*dx*dt/dA (main.cpp:246-253) as the depth added.
    static __device__ __forceinline__ T source_depth(const ScnDev& S, const TileCtx<T>& x, int k) {
        const int off = glob<const int>(S.src_off)[k];
        const double val = interp_series_short(glob<const double>(S.series_t) + off,
                                               glob<const double>(S.series_v) + off,
                                               glob<const int>(S.src_len)[k],
                                               glob<const double>(S.src_val)[k], x.t);
        const double ps = val * S.dx * x.dt / x.dA;
        return (T)ps;
    }

    // Value of override entry k: 0 (FREE), the fixed depth (HFIX) or the
    // depth series at the step's start time (HVAR); kernels.cpp:266-273,
    // main.cpp:275-279.
    static __device__ __forceinline__ T override_depth(const ScnDev& S, const TileCtx<T>& x,
                                                       int k) {
        const int off = glob<const int>(S.ovr_off)[k];
        const double val = interp_series_short(glob<const double>(S.series_t) + off,
                                               glob<const double>(S.series_v) + off,
                                               glob<const int>(S.ovr_len)[k],
                                               glob<const double>(S.ovr_val)[k], x.t);
        return (T)val;
    }

    // At the tile's start: the first kPointBatch sources and overrides of the
    // tile, evaluated across the lanes (entry base + 64 r + lane in round r)
    // into the wave's LDS slice.  Entries past the batch are taken one at a
    // time in the row loop.
    static __device__ __forceinline__ void load_points(const ScnDev& S, const TileCtx<T>& x,
                                                       TileState<T, V>& st) {
        st.sn = min(kPointBatch, st.se - st.sp);
        st.on = min(kPointBatch, st.oe - st.op);
        for (int r = 0; r < st.sn; r += 64) {
            const int e = r + x.lane;
            if (e < st.sn) {
                const int k = st.sp + e;
                const int j = glob<const int>(S.src_j)[k];
                const int col = glob<const int>(S.src_i)[k];
                x.lrc[e] = ((j - x.j0) << 9) | (col - x.i0);
                x.lval[e] = source_depth(S, x, k);
            }
        }
        for (int r = 0; r < st.on; r += 64) {
            const int e = r + x.lane;
            if (e < st.on) {
                const int k = st.op + e;
                const int j = glob<const int>(S.ovr_j)[k];
                const int col = glob<const int>(S.ovr_i)[k];
                x.lrc[kPointBatch + e] = ((j - x.j0) << 9) | (col - x.i0);
                x.lval[kPointBatch + e] = override_depth(S, x, k);
            }
        }
        st.sk = 0;
        st.ok = 0;
        st.src_next_j = (st.sn > 0) ? x.j0 + (__builtin_amdgcn_readfirstlane(x.lrc[0]) >> 9)
                                    : INT_MAX;
        st.ovr_next_j = (st.on > 0)
                            ? x.j0 + (__builtin_amdgcn_readfirstlane(x.lrc[kPointBatch]) >> 9)
                            : INT_MAX;
    }

    // The next source, on row src_next_j: its depth is added to its cell
    // with a clamp (main.cpp:256-260), in list order, which is file order on
    // a shared cell.
    static __device__ __forceinline__ void apply_source(const ScnDev& S, const TileCtx<T>& x,
                                                        TileState<T, V>& st, T (&hn)[V]) {
        const int k = st.sk;
        int col;
        T dh;
        if (k < st.sn) {
            col = x.i0 + (__builtin_amdgcn_readfirstlane(x.lrc[k]) & 511);
            dh = x.lval[k];
        } else {  // past the batch: one entry at a time, uniform loads
            col = __builtin_amdgcn_readfirstlane(glob<const int>(S.src_i)[st.sp + k]);
            dh = source_depth(S, x, st.sp + k);
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (col == x.c0 + v) {
                T h = hn[v] + dh;
                if (h < T(0)) h = T(0);
                hn[v] = h;
            }
        }
        const int n = k + 1;
        st.sk = n;
        if (n < st.sn)
            st.src_next_j = x.j0 + (__builtin_amdgcn_readfirstlane(x.lrc[n]) >> 9);
        else if (st.sp + n < st.se)
            st.src_next_j = __builtin_amdgcn_readfirstlane(glob<const int>(S.src_j)[st.sp + n]);
        else
            st.src_next_j = INT_MAX;
    }

    // The next FREE/HFIX/HVAR point: the last in file order on a cell wins.
    static __device__ __forceinline__ void apply_override(const ScnDev& S, const TileCtx<T>& x,
                                                          TileState<T, V>& st, T (&hn)[V]) {
        const int k = st.ok;
        int col;
        T val;
        if (k < st.on) {
            col = x.i0 + (__builtin_amdgcn_readfirstlane(x.lrc[kPointBatch + k]) & 511);
            val = x.lval[kPointBatch + k];
        } else {
            col = __builtin_amdgcn_readfirstlane(glob<const int>(S.ovr_i)[st.op + k]);
            val = override_depth(S, x, st.op + k);
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (col == x.c0 + v) hn[v] = val;
        }
        const int n = k + 1;
        st.ok = n;
        if (n < st.on)
            st.ovr_next_j = x.j0 + (__builtin_amdgcn_readfirstlane(x.lrc[kPointBatch + n]) >> 9);
        else if (st.op + n < st.oe)
            st.ovr_next_j = __builtin_amdgcn_readfirstlane(glob<const int>(S.ovr_j)[st.op + n]);
        else
            st.ovr_next_j = INT_MAX;
    }

    // EDGE = false: the strip touches neither the west nor the east edge and
    // every lane's columns are cells, so no per-face edge test is needed.
    template <bool EDGE>
    static __device__ __forceinline__ void run(const ScnDev& S, const TileDesc td, int kmod6, int lane,
                                               lptr<int> lrc, lptr<T> lval) {
        const int cur = kmod6 & 1, nxt = cur ^ 1;
        const int hs_cur = kmod6 % 3, hs_nxt = (kmod6 + 1) % 3, hs_rst = (kmod6 + 2) % 3;
        gptr<Ctl> ctl = glob<Ctl>(S.ctl);
        const int tile_local = td.strip + td.seg * S.nstrips;

        // state buffers: a skipped step leaves the state where it is, so the
        // buffer pair follows the scenario's own executed steps
        const int s_in = ctl->par[cur], s_out = s_in ^ 1;
        const double t = ctl->t[cur];
        const double next_mass = ctl->next_mass[cur];
        const double hmax_in = fkey_inv_d(ctl->hmax[hs_cur]);
        const int status = ctl->status[cur];
        const double dt = cfl_dt(hmax_in, S.thresh, S.cfl, S.dx, S.g, S.init_tstep);
        // a step that cannot move the clock (dt lost against t) would repeat
        // forever at the same t: every wave sees the same t and dt, so all of
        // them skip it, and the first tile's wave marks the scenario stalled
        const bool running = (t < S.sim_time) && status == 0;  // main.cpp:227
        const bool stalled = running && !(t + dt > t);
        const bool skip = !running || stalled;
        const bool report_due = !skip && (t + dt >= next_mass);

        if (td.first) bookkeeping(S, cur, nxt, hs_cur, hs_nxt, hs_rst, s_in, t, dt, next_mass,
                                  skip, report_due, lane);
        if (td.first && stalled && lane == 0) atomicOr(&S.ctl->status[nxt], kStatusStalled);
        if (skip) return;

        TileCtx<T> x;
        x.nx = S.nx;
        x.ny = S.ny;
        x.P = S.P;
        x.i0 = td.strip * W;
        const int j0 = td.seg * S.tile_rows;
        x.j0 = j0;
        x.j1 = min(j0 + S.tile_rows, x.ny);
        x.c0 = x.i0 + lane * V;
        x.lane = lane;
        x.lrc = lrc;
        x.lval = lval;
        if constexpr (F64) x.nod = S.nodata_t;
        else x.nod = INT_MIN;
        x.znod = (T)(S.nodata_t - S.zbase);
        x.z = glob<const Zt<T>>(S.z);
        x.Ho = glob<const T>(S.H[s_in]);
        x.Hn = glob<T>(S.H[s_out]);
        x.Qxo = glob<const T>(S.Qx[s_in]);
        x.Qxn = glob<T>(S.Qx[s_out]);
        x.Qyo = glob<const T>(S.Qy[s_in]);
        x.Qyn = glob<T>(S.Qy[s_out]);
        x.sc = (T)S.dxinv;  // Q -> q_old (kernels.cpp:8-16)
        x.dxT = (T)S.dx;
        x.gT = (T)S.g;
        x.kd = FluxConstsF64{S.dx, S.fn, S.thresh, S.maxh, S.g * dt, S.theta};
        x.kf = FluxConstsF32{(float)S.dx, (float)S.thresh, (float)S.maxh,
                             (float)(S.g * dt / S.dx), (float)(S.g * dt * S.fn * S.fn),
                             (float)S.dxinv, (float)(0.25 * S.dxinv), (float)S.zscale};
        x.t = t;
        x.dt = dt;
        x.dA = S.dA;
        x.dtdA = (T)(dt / S.dA);
        x.report_due = report_due;
        // rain depth of this step (main.cpp:265-268, kernels.cpp:356-358)
        double rain_depth = 0.0;
        x.rain_on = false;
        if (S.rain_enabled) {
            double r = interp_series_short(glob<const double>(S.rain_t),
                                           glob<const double>(S.rain_r), S.n_rain, 0.0, t);
            x.rain_on = (r > 0.0);
            if (x.rain_on) rain_depth = r * dt;
        }
        x.rainT = (T)rain_depth;
        // halo columns: lane 0 reads column i0-1, every other lane column i0+W
        // (only lane 63 uses it; the others hit the same cache line)
        const int hcol = (lane == 0) ? x.i0 - 1 : x.i0 + W;
        x.hcell = !EDGE || ((lane == 0) ? (x.i0 > 0) : (x.i0 + W < x.nx));  // a cell column
        x.hface = (lane != 0) && (!EDGE || x.i0 + W <= x.nx);              // an x-face column
        // a halo column outside the grid is loaded from the wave's own first
        // column instead: that line is fetched anyway, and the load keeps the
        // wait counts the same on every path
        x.hcol_c = x.hcell ? hcol : x.i0;
        x.hcol_f = x.hface ? hcol : x.i0;
        // rows are needed up to j1 (the south neighbour of the tile's last row)
        x.rc = min(x.ny, x.j1 + 1);
        x.rf = min(x.ny + 1, x.j1 + 1);

        TileState<T, V> st;
        RowData<T, V> A, B, C, D;
        load(A, x, j0);
        load(B, x, j0 + 1);

        // north faces of row j0
        if (j0 == 0) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                T q = T(0);
                if (S.north_free && (!EDGE || x.c0 + v < x.nx))
                    q = edge_out_neg<T>(A.h[v], A.z[v] == x.nod, x.dxT, x.gT);
                st.qyN[v] = q;
            }
            if (!EDGE || x.c0 < x.P) {
                Vec<T, V> o;
#pragma unroll
                for (int v = 0; v < V; v++) o[v] = (!EDGE || x.c0 + v < x.nx) ? st.qyN[v] : T(0);
                *(gptr<Vec<T, V>>)(x.Qyn + x.c0) = o;
            }
        } else {
            T ha[V], qxa[V];
            Zt<T> za[V];
            load_row<T, V>(ha, x.Ho, x.P, j0 - 1, x.ny, x.c0);
            load_row<Zt<T>, V>(za, x.z, x.P, j0 - 1, x.ny, x.c0);
            load_row<T, V>(qxa, x.Qxo, x.P, j0 - 1, x.ny, x.c0);
            T qxH_a = load_one(x.Qxo, x.P, j0 - 1, x.ny, x.hcol_f);
            T qxa_n = lane_next(qxa[0], qxH_a, lane);
            T qxc_n = lane_next(A.qx[0], A.qxH, lane);
#pragma unroll
            for (int v = 0; v < V; v++) {
                T a1 = (v + 1 < V) ? qxa[v + 1] : qxa_n;
                T c1 = (v + 1 < V) ? A.qx[v + 1] : qxc_n;
                st.qyN[v] = (!EDGE || x.c0 + v < x.nx)
                                ? flux(za[v], A.z[v], ha[v], A.h[v], A.qy[v], qxa[v], a1, A.qx[v], c1,
                                       x.nod, x.sc, x.kd, x.kf)
                                : T(0);
            }
        }

        // per-tile point lists (uniform)
        gptr<const int> src_begin = glob<const int>(S.src_begin);
        gptr<const int> ovr_begin = glob<const int>(S.ovr_begin);
        st.sp = __builtin_amdgcn_readfirstlane(src_begin[tile_local]);
        st.se = __builtin_amdgcn_readfirstlane(src_begin[tile_local + 1]);
        st.op = __builtin_amdgcn_readfirstlane(ovr_begin[tile_local]);
        st.oe = __builtin_amdgcn_readfirstlane(ovr_begin[tile_local + 1]);
        st.sn = st.on = st.sk = st.ok = 0;
        st.src_next_j = INT_MAX;
        st.ovr_next_j = INT_MAX;
        if (st.sp < st.se || st.op < st.oe) load_points(S, x, st);
        st.hmax = F64 ? T(-DBL_MAX) : T(-FLT_MAX);
        st.finite = true;
        st.vol_s = 0.0;
        st.vol_c = 0.0;
        st.wet = 0;

        const int j1 = x.j1;
        if constexpr (PF == 1) {
            // three register sets rotate through the roles (row j, row j+1,
            // row j+2 in flight); unrolling by three keeps each in fixed registers
#pragma unroll 1
            for (int j = j0; j < j1; j += 3) {
                row<EDGE>(S, x, st, A, B, C, j);
                if (j + 1 >= j1) break;
                row<EDGE>(S, x, st, B, C, A, j + 1);
                if (j + 2 >= j1) break;
                row<EDGE>(S, x, st, C, A, B, j + 2);
            }
        } else {
            // two rows per step of the loop; four register sets alternate
            // between resident and in flight
#pragma unroll 1
            for (int j = j0; j < j1; j += 4) {
                row2<EDGE>(S, x, st, A, B, C, D, j);
                if (j + 2 >= j1) break;
                row2<EDGE>(S, x, st, C, D, A, B, j + 2);
            }
        }

        // a non-finite depth stops the scenario from the next step on (the
        // bit is sticky; steps already under way finish or skip, and the
        // scenario's results are not used)
        if (__any(!st.finite) && lane == 0) atomicOr(&S.ctl->status[nxt], kStatusNonfinite);

        // wavefront max, one atomic per wavefront
        T hm = st.hmax;
        for (int off = 32; off > 0; off >>= 1) hm = getmax(hm, __shfl_xor(hm, off, 64));
        if (lane == 0) {
            // the empty-set identity stays the lowest double, as a Kokkos max reducer's
            const double m = (hm == (F64 ? T(-DBL_MAX) : T(-FLT_MAX))) ? -DBL_MAX : (double)hm;
            atomicMax(&S.ctl->hmax[hs_nxt], fkey(m));
        }
        if (report_due) {
            double vol_s = st.vol_s, vol_c = st.vol_c;
            unsigned long long wet = st.wet;
            for (int off = 32; off > 0; off >>= 1) {
                double os = __shfl_xor(vol_s, off, 64);
                double oc = __shfl_xor(vol_c, off, 64);
                unsigned long long ow = __shfl_xor(wet, off, 64);
                // lanes pair up symmetrically; both keep the same merged pair
                double a = (lane & off) ? os : vol_s;
                double b = (lane & off) ? vol_s : os;
                double ca = (lane & off) ? oc : vol_c;
                double cb = (lane & off) ? vol_c : oc;
                double s2 = a, c2 = ca + cb;
                kahan_add(s2, c2, b);
                vol_s = s2;
                vol_c = c2;
                wet += ow;
            }
            if (lane == 0) {
                const long long slot = (long long)cur * S.ntiles + tile_local;
                glob<double>(S.part_vol)[slot] = vol_s;
                glob<double>(S.part_comp)[slot] = vol_c;
                glob<unsigned long long>(S.part_area)[slot] = wet;
            }
        }
    }

    // Scenario bookkeeping, done by the wavefront of the scenario's first tile:
    // finalise the previous step's mass report, advance t, rotate the H-max
    // slots, log dt.
    static __device__ __forceinline__ void bookkeeping(const ScnDev& S, int cur, int nxt, int hs_cur, int hs_nxt,
                                       int hs_rst, int s_in, double t, double dt,
                                       double next_mass, bool skip, bool report_due, int lane) {
        gptr<Ctl> ctl = glob<Ctl>(S.ctl);
        finalize_report(S, lane);
        if (lane == 0) {
            // the status bits step k+1 reads: those of step k, and any that
            // the waves of step k set (atomically) into the same word
            atomicOr(&S.ctl->status[nxt], ctl->status[cur]);
            ctl->hmax[hs_rst] = key_neg_max();  // the slot step k+1 max-reduces into
            if (!skip) {
                const double t_new = t + dt;  // main.cpp:287
                ctl->t[nxt] = t_new;
                ctl->par[nxt] = s_in ^ 1;
                const long long n = ctl->n_done;
                if (n < S.dt_log_cap) glob<double>(S.dt_log)[n] = dt;
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
                ctl->par[nxt] = s_in;
                ctl->next_mass[nxt] = next_mass;
                ctl->hmax[hs_nxt] = ctl->hmax[hs_cur];
            }
        }
    }

    static __device__ __forceinline__ void finalize_report(const ScnDev& S, int lane) {
        gptr<Ctl> ctl = glob<Ctl>(S.ctl);
        if (!ctl->pending) return;
        const long long base = (long long)ctl->pending_buf * S.ntiles;
        double s = 0.0, c = 0.0;
        unsigned long long wet = 0;
        for (int i = lane; i < S.ntiles; i += 64) {
            kahan_add(s, c, glob<const double>(S.part_vol)[base + i]);
            c += glob<const double>(S.part_comp)[base + i];
            wet += glob<const unsigned long long>(S.part_area)[base + i];
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
                gptr<double> rec = glob<double>(S.reports) + 4LL * r;
                rec[0] = ctl->pending_t;
                rec[1] = ctl->pending_dt;
                rec[2] = s + c;
                rec[3] = (double)wet * S.dA;
            }
            ctl->n_reports = r + 1;
            ctl->last_wet = wet;
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

// Static LDS of a k_step workgroup: each wave's batched point entries.
template <typename T, int WPB>
constexpr size_t step_static_lds() {
    return (size_t)WPB * 2 * kPointBatch * (sizeof(int) + sizeof(T));
}

template <typename T, int V, bool F64, int WPB, int OCC, int PF>
__global__ void __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(OCC)))
k_step(const ScnDev* __restrict__ scns, const TileDesc* __restrict__ tiles, int ntiles, int kmod6) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tile = lb * WPB + wv;
    if (tile >= ntiles) return;
    const TileDesc td = tiles[tile];
    const ScnDev& S = scns[td.scn];
    constexpr int W = 64 * V;
    // per wave: the batched source and override entries of the tile
    __shared__ int l_rc[WPB][2 * kPointBatch];
    __shared__ T l_val[WPB][2 * kPointBatch];
    static_assert(sizeof(l_rc) + sizeof(l_val) == step_static_lds<T, WPB>(), "static LDS");
    const lptr<int> lrc = (lptr<int>)&l_rc[wv][0];
    const lptr<T> lval = (lptr<T>)&l_val[wv][0];
    if (td.strip > 0 && (td.strip + 1) * W < S.nx)
        Stepper<T, V, F64, PF>::template run<false>(S, td, kmod6, lane, lrc, lval);
    else
        Stepper<T, V, F64, PF>::template run<true>(S, td, kmod6, lane, lrc, lval);
}

}  // namespace af
