// fused.hip -- host runtime of the fused step: arena layout, point plans,
// uploads, eager and graph-replayed stepping, read-back, and the C ABI.
#include <float.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "af_host.hpp"
#include "fused_kernel.hpp"

namespace af {
namespace {

inline unsigned long long host_key(double x) {
    unsigned long long b;
    memcpy(&b, &x, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bump layout of one arena: offsets first, pointers after the arena is known.
struct Layout {
    size_t top = 0;
    size_t take(size_t bytes) {
        size_t o = align_up(top, 256);
        top = o + align_up(bytes ? bytes : 1, 256);
        return o;
    }
};

struct HostScn {
    int nx = 0, ny = 0;
    long long P = 0;
    int nstrips = 0, nsegs = 0, ntiles = 0, tile_base = 0;
    int R = 0;                   // tile rows of this scenario
    int R_alone = 0;             // tile rows it would take running alone (<= R)
    int cap_tiles = 0;           // tiles at R_alone: the per-tile arrays are sized for it
    double zbase = 0.0, zscale = 1.0;   // fixed-point terrain of the FP32 path
    long long wet0 = 0;          // cells deeper than 0.01 m in the initial state
    // offsets in the arena
    size_t o_z, o_H[2], o_Qx[2], o_Qy[2];
    size_t o_src_begin, o_src_j, o_src_i, o_src_val, o_src_off, o_src_len;
    size_t o_ovr_begin, o_ovr_j, o_ovr_i, o_ovr_val, o_ovr_off, o_ovr_len;
    size_t o_series_t, o_series_v;
    size_t o_rain_t, o_rain_r, o_ctl, o_part_vol, o_part_comp, o_part_area, o_dt_log, o_reports;
    // host-built plan contents
    std::vector<int> src_begin, src_j, src_i, src_off, src_len;
    std::vector<int> ovr_begin, ovr_j, ovr_i, ovr_off, ovr_len;
    std::vector<double> src_val, ovr_val;
    int n_series = 0, n_rain = 0;
    ScnDev dev{};
    // pool bookkeeping
    bool active = false;
    size_t base = 0, bytes = 0;  // heap extent of this scenario in the arena
    int base_parity = 0;         // state buffer holding the initial state
    bool in_lane = false;        // stepped by the lane (af_pool_set_lane), not the main table
};

// Host staging memory: pinned (page-locked) blocks, so that the copies to the
// device are plain DMA transfers that do not pass through the runtime's
// pageable staging.  Blocks come in power-of-two sizes from 64 KiB and go
// back to a process-wide cache when released, never to the system:
// pinning is slow, and hipHostFree can synchronise the device.  The cache
// therefore holds the peak of staging memory in use at once.  Where pinned
// memory cannot be had, a block falls back to ordinary memory.
class HostBuf {
  public:
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void* data() const { return p_; }
    bool reserve(size_t bytes) {
        release();
        size_t cls = size_t(1) << 16;
        while (cls < bytes) cls <<= 1;
        {
            std::lock_guard<std::mutex> g(mu());
            auto it = cache().find(cls);
            if (it != cache().end()) {
                p_ = it->second;
                cache().erase(it);
                cls_ = cls;
                pinned_ = true;
                return true;
            }
        }
        if (hipHostMalloc(&p_, cls, hipHostMallocPortable) == hipSuccess) {
            pinned_ = true;
        } else {
            (void)hipGetLastError();
            p_ = malloc(cls);
            pinned_ = false;
            if (!p_) return false;
        }
        cls_ = cls;
        return true;
    }
    void release() {
        if (!p_) return;
        if (pinned_) {
            std::lock_guard<std::mutex> g(mu());
            cache().emplace(cls_, p_);
        } else {
            free(p_);
        }
        p_ = nullptr;
    }

  private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::multimap<size_t, void*>& cache() {
        static std::multimap<size_t, void*> c;
        return c;
    }
    void* p_ = nullptr;
    size_t cls_ = 0;
    bool pinned_ = false;
};

// A scenario's terrain and initial depths in the device format, converted on
// the host before admission (stage_inputs): the part of an admission that
// needs neither the pool nor the GPU, so it can run while the pool steps.
struct Staged {
    int precision = AF_FP32;
    int nx = 0, ny = 0;
    long long P = 0;
    HostBuf z, h;   // P*ny elements each: FP64 elevations and depths, or the
                    // FP32 path's fixed-point terrain (int) and float depths
    double zbase = 0.0, zscale = 1.0, hmax = -DBL_MAX;
    long long wet = 0;   // cells deeper than 0.01 m (the Area criterion of main.cpp:296)
};

}  // namespace
}  // namespace af

// A scenario staged by af_pool_stage: a copy of its descriptor (whose arrays
// the caller keeps) and its converted terrain and depths.
struct af_staged {
    af_scenario_desc desc{};
    af::Staged s;
};

struct af_domain {
    int precision = AF_FP32;
    int V = 4, W = 256, R = 128, wpb = 4, graph_steps = 36, variant = 0;
    int R_fixed = 0;             // tile rows asked for (0: chosen per scenario)
    int tile_order = 0;          // af_tuning.tile_order
    int min_rows = 8;            // lowest default tile height (see kMinRowsAlone)
    bool pool = false;           // created by af_pool_create
    long long dt_log_cap = 0;
    int report_cap = 64;
    char* arena = nullptr;
    size_t arena_bytes = 0;
    // pool: a fixed table region, then a heap of scenario extents
    int max_scn = 0, max_tiles = 0;
    std::vector<af::HostScn> scn;  // max_scn slots
    std::vector<std::pair<size_t, size_t>> free_list;  // (offset, size), sorted
    size_t o_scns = 0, o_tiles = 0, o_volbuf = 0, o_ctlptr = 0, o_ctlbuf = 0, heap_lo = 0;
    size_t o_lane = 0;
    af::ScnDev* d_scns = nullptr;
    af::TileDesc* d_tiles = nullptr;
    double* d_volbuf = nullptr;
    af::Ctl** d_ctlptr = nullptr;     // control blocks of the active slots, in `active` order
    af::Ctl* d_ctlbuf = nullptr;      // their gathered copies
    std::vector<int> active;          // active slots, ascending (as of the last tile rebuild)
    int ntiles = 0;
    bool dirty = false;               // admissions or releases since the last tile rebuild
    long long k = 0;  // index of the next step to launch
    double last_tsum = -1.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // the lane: a second tile table holding every tile of some scenarios
    // (af_pool_set_lane), which the main table then leaves out;
    // af_pool_advance_lane steps them on a stream of its own, concurrently
    // with the main table's launches, until a call that joins it
    std::vector<int> lane_slots;
    af::TileDesc* d_lane = nullptr;
    int nlane = 0;
    hipGraph_t lgraph = nullptr;
    hipGraphExec_t lgexec = nullptr;
    hipStream_t lane_st = nullptr;
    hipEvent_t ev_fork = nullptr;     // the caller's stream when the lane work was enqueued
    hipEvent_t ev_t0 = nullptr;       // the lane work's start and end (timed)
    hipEvent_t ev_join = nullptr;
    bool lane_pending = false;        // lane work not yet joined into the caller's stream
    bool lane_timed = false;          // ev_t0 / ev_join hold a lane burst
    int lds_cu = 0, lds_block = 0;    // the device's LDS bytes per CU and per workgroup
    // FP64 with af_tuning.variant 0: the kernel variant follows the wet area
    // (adapt_variant); most_tiles is the largest scenario's tile count
    bool variant_auto = false;
    int most_tiles = 0;
};

namespace af {
namespace {

constexpr int kVolBlocks = 1024;
constexpr int kMinTiles = 2048;   // wavefronts a lone scenario should offer
// Lowest default tile height: a scenario that runs by itself (a lone domain,
// or one admitted to a pool holding fewer than kMinTiles tiles) goes down to
// kMinRowsAlone rows; one that shares the launch with enough other tiles
// stops at kMinRowsShared, where the rows re-read above and below each
// segment cost a quarter more reads instead of half or more.
constexpr int kMinRowsAlone = 2;
constexpr int kMinRowsShared = 8;
// Tiles shorter than 8 rows are taken only while the scenario stays under
// kMinTiles tiles, and shorter than 4 only while it stays at or under
// kMaxTilesTiny: a lone domain's step is bound by the instructions each SIMD
// issues, and past that many tiles the per-tile work (north faces, point
// lists, the first rows' loads) outweighs the shorter row march.
constexpr int kMaxTilesTiny = 512;

int elem_size(int precision) { return precision == AF_FP64 ? 8 : 4; }

int validate(const af_scenario_desc* d) {
    AF_REQUIRE(d, AF_E_INVALID, "null scenario");
    AF_REQUIRE(d->nx >= 2 && d->ny >= 2, AF_E_INVALID, "grid %d x %d: need at least 2 x 2", d->nx,
               d->ny);
    AF_REQUIRE((long long)d->nx * d->ny < (1LL << 31), AF_E_INVALID,
               "grid %d x %d: more than 2^31 cells", d->nx, d->ny);
    AF_REQUIRE(d->dem, AF_E_INVALID, "null DEM");
    AF_REQUIRE(d->dx > 0.0 && std::isfinite(d->dx), AF_E_INVALID,
               "cell size must be positive and finite");
    // a NaN or infinite parameter would reach every cell on the first step;
    // it is refused here instead (the reference runs it and returns NaN)
    const struct { const char* name; double v; } params[] = {
        {"fpn", d->fpn}, {"depth_thresh", d->depth_thresh}, {"max_hflow", d->max_hflow},
        {"init_tstep", d->init_tstep}, {"cfl", d->cfl}, {"g", d->g},
        {"sim_time", d->sim_time}, {"mass_int", d->mass_int}};
    for (const auto& p : params)
        AF_REQUIRE(std::isfinite(p.v), AF_E_INVALID, "%s is not finite (%g)", p.name, p.v);
    AF_REQUIRE(d->nodata == d->nodata, AF_E_INVALID, "NODATA value is NaN");
    AF_REQUIRE(d->theta == 1.0, AF_E_INVALID,
               "the fused step implements theta = 1 (main.cpp:124); got %g", d->theta);
    AF_REQUIRE(d->n_points >= 0, AF_E_INVALID, "negative point count");
    if (d->n_points > 0)
        AF_REQUIRE(d->pt_x && d->pt_y && d->pt_type && d->pt_value, AF_E_INVALID,
                   "null point arrays");
    for (int p = 0; p < d->n_points; p++) {
        long long idx = (long long)d->pt_x[p] + (long long)d->pt_y[p] * d->nx;  // io.cpp:187
        AF_REQUIRE(idx >= 0 && idx < (long long)d->nx * d->ny, AF_E_INVALID,
                   "point %d at grid (%d, %d) lies outside the %d x %d grid", p, d->pt_x[p],
                   d->pt_y[p], d->nx, d->ny);
        int ty = d->pt_type[p];
        AF_REQUIRE(ty >= AF_BC_CLOSED && ty <= AF_BC_QVAR, AF_E_INVALID, "point %d: type %d", p,
                   ty);
        int len = d->pt_series_len ? d->pt_series_len[p] : 0;
        int off = d->pt_series_off ? d->pt_series_off[p] : 0;
        AF_REQUIRE(len >= 0 && off >= 0 && (long long)off + len <= d->n_series, AF_E_INVALID,
                   "point %d: series [%d, %d) outside %d entries", p, off, off + len,
                   d->n_series);
        if (len > 0) AF_REQUIRE(d->series_t && d->series_v, AF_E_INVALID, "null series arrays");
        // a FREE point's value is not used (kernels.cpp:266-268)
        AF_REQUIRE(ty == AF_BC_FREE || ty == AF_BC_CLOSED || std::isfinite(d->pt_value[p]),
                   AF_E_INVALID, "point %d: value %g is not finite", p, d->pt_value[p]);
        for (int e = off; e < off + len; e++)
            AF_REQUIRE(std::isfinite(d->series_t[e]) && std::isfinite(d->series_v[e]),
                       AF_E_INVALID, "point %d: series entry %d is not finite (%g at t = %g)", p,
                       e - off, d->series_v[e], d->series_t[e]);
    }
    if (d->rain_enabled) {
        AF_REQUIRE(d->n_rain > 0 && d->rain_t && d->rain_r, AF_E_INVALID,
                   "rain enabled without a series");
        for (int e = 0; e < d->n_rain; e++)
            AF_REQUIRE(std::isfinite(d->rain_t[e]) && std::isfinite(d->rain_r[e]), AF_E_INVALID,
                       "rain entry %d is not finite (%g at t = %g)", e, d->rain_r[e],
                       d->rain_t[e]);
    }
    return AF_OK;
}

// Adds the heap base of the scenario's extent to every offset.
void rebase(HostScn& h, size_t b) {
    size_t* fields[] = {&h.o_z, &h.o_H[0], &h.o_H[1], &h.o_Qx[0], &h.o_Qx[1], &h.o_Qy[0],
                        &h.o_Qy[1], &h.o_src_begin, &h.o_src_j, &h.o_src_i, &h.o_src_val,
                        &h.o_src_off, &h.o_src_len, &h.o_ovr_begin, &h.o_ovr_j, &h.o_ovr_i,
                        &h.o_ovr_val, &h.o_ovr_off, &h.o_ovr_len, &h.o_series_t, &h.o_series_v,
                        &h.o_rain_t, &h.o_rain_r, &h.o_ctl, &h.o_part_vol, &h.o_part_comp,
                        &h.o_part_area, &h.o_dt_log, &h.o_reports};
    for (size_t* f : fields) *f += b;
    h.base = b;
}

// Row pitch in elements: wide rows start on an odd multiple of 1 KiB, which
// spreads the rows a strip marches down over the HBM channels (a
// 256-byte-aligned pitch of nx+1 elements measured 7% slower on the
// 4-read/3-write row march at 16384 columns); narrow rows keep 256-byte
// alignment.
long long row_pitch(int nx, int precision) {
    const size_t es = elem_size(precision);
    size_t row = ((size_t)nx + 1) * es;
    if (row >= 8192) {
        row = align_up(row, 1024);
        if (((row >> 10) & 1) == 0) row += 1024;
    } else {
        row = align_up(row, 256);
    }
    return (long long)(row / es);
}

void plan_geometry(HostScn& h, const af_scenario_desc* d, const af_domain* D) {
    h.nx = d->nx;
    h.ny = d->ny;
    h.P = row_pitch(d->nx, D->precision);
    h.nstrips = (d->nx + D->W - 1) / D->W;
    // Tile rows: the tuning's value, else the default height, halved while
    // the scenario alone would give fewer than kMinTiles wavefronts, down to
    // D->min_rows (below 8 and 4 rows only while the tiles stay few; see
    // kMaxTilesTiny).  A small domain run by itself is latency-bound: each
    // wavefront marches its rows one after another, so a step takes as long
    // as the tallest tile, and shorter tiles split the rows over more
    // wavefronts.
    auto tiles = [&](int r) { return (long long)h.nstrips * ((d->ny + r - 1) / r); };
    auto halve = [&](int floor_rows) {
        int r = D->R;
        while (r > floor_rows && tiles(r) < kMinTiles) {
            const int half = r / 2;
            if (half < 8 && tiles(half) >= kMinTiles) break;
            if (half < 4 && tiles(half) > kMaxTilesTiny) break;
            r = half;
        }
        return r;
    };
    int R = D->R_fixed, Ra = D->R_fixed;
    if (R <= 0) {
        R = halve(D->min_rows);
        Ra = halve(kMinRowsAlone);
    }
    h.R = R;
    h.R_alone = Ra;
    h.nsegs = (d->ny + R - 1) / R;
    h.ntiles = h.nstrips * h.nsegs;
    h.cap_tiles = h.nstrips * ((d->ny + Ra - 1) / Ra);
}

void plan_layout(HostScn& h, const af_scenario_desc* d, const af_domain* D, Layout& L) {
    const size_t es = elem_size(D->precision);
    const size_t cells = (size_t)h.P * h.ny, faces_y = (size_t)h.P * (h.ny + 1);
    h.o_z = L.take(cells * es);
    for (int b = 0; b < 2; b++) {
        h.o_H[b] = L.take(cells * es);
        h.o_Qx[b] = L.take(cells * es);
        h.o_Qy[b] = L.take(faces_y * es);
    }
    int n_src = 0, n_ovr = 0;
    for (int p = 0; p < d->n_points; p++) {
        int ty = d->pt_type[p];
        if (ty == AF_BC_QFIX || ty == AF_BC_QVAR) n_src++;
        else if (ty == AF_BC_FREE || ty == AF_BC_HFIX || ty == AF_BC_HVAR) n_ovr++;
    }
    const size_t nt1 = (size_t)h.cap_tiles + 1;
    h.o_src_begin = L.take(nt1 * 4);
    h.o_src_j = L.take(n_src * 4);
    h.o_src_i = L.take(n_src * 4);
    h.o_src_val = L.take(n_src * 8);
    h.o_src_off = L.take(n_src * 4);
    h.o_src_len = L.take(n_src * 4);
    h.o_ovr_begin = L.take(nt1 * 4);
    h.o_ovr_j = L.take(n_ovr * 4);
    h.o_ovr_i = L.take(n_ovr * 4);
    h.o_ovr_val = L.take(n_ovr * 8);
    h.o_ovr_off = L.take(n_ovr * 4);
    h.o_ovr_len = L.take(n_ovr * 4);
    h.o_series_t = L.take((size_t)d->n_series * 8);
    h.o_series_v = L.take((size_t)d->n_series * 8);
    h.o_rain_t = L.take((size_t)(d->rain_enabled ? d->n_rain : 0) * 8);
    h.o_rain_r = L.take((size_t)(d->rain_enabled ? d->n_rain : 0) * 8);
    h.o_ctl = L.take(sizeof(Ctl));
    h.o_part_vol = L.take(2 * (size_t)h.cap_tiles * 8);
    h.o_part_comp = L.take(2 * (size_t)h.cap_tiles * 8);
    h.o_part_area = L.take(2 * (size_t)h.cap_tiles * 8);
    h.o_dt_log = L.take((size_t)D->dt_log_cap * 8);
    h.o_reports = L.take((size_t)D->report_cap * 32);
}

// Points sorted per tile by (row, column, file order): a stable sort keeps
// file order among points of one cell, which is the order the reference's
// sequential loops would apply them in.  Each entry carries what its value
// needs: a constant (QFIX and HFIX value, 0 for FREE) with an empty series,
// or a series (QVAR, HVAR) whose empty-series value is the point's value.
// Interpolating an empty series returns that value, so every entry is
// evaluated the same way, from the entry itself (main.cpp:245-253, 275-279).
void plan_points(HostScn& h, const af_scenario_desc* d, const af_domain* D) {
    struct Item { int tile, j, i, pt; };
    std::vector<Item> src, ovr;
    for (int p = 0; p < d->n_points; p++) {
        long long idx = (long long)d->pt_x[p] + (long long)d->pt_y[p] * d->nx;
        int j = (int)(idx / d->nx), i = (int)(idx % d->nx);
        int tile = (i / D->W) + (j / h.R) * h.nstrips;
        int ty = d->pt_type[p];
        if (ty == AF_BC_QFIX || ty == AF_BC_QVAR) src.push_back({tile, j, i, p});
        else if (ty == AF_BC_FREE || ty == AF_BC_HFIX || ty == AF_BC_HVAR) ovr.push_back({tile, j, i, p});
    }
    auto key_less = [](const Item& a, const Item& b) {
        if (a.tile != b.tile) return a.tile < b.tile;
        if (a.j != b.j) return a.j < b.j;
        return a.i < b.i;
    };
    std::stable_sort(src.begin(), src.end(), key_less);
    std::stable_sort(ovr.begin(), ovr.end(), key_less);
    auto build = [&](const std::vector<Item>& v, std::vector<int>& begin, std::vector<int>& J,
                     std::vector<int>& I, std::vector<double>& VAL, std::vector<int>& OFF,
                     std::vector<int>& LEN) {
        begin.assign(h.ntiles + 1, 0);
        for (const Item& it : v) begin[it.tile + 1]++;
        for (int t = 0; t < h.ntiles; t++) begin[t + 1] += begin[t];
        J.clear(); I.clear(); VAL.clear(); OFF.clear(); LEN.clear();
        for (const Item& it : v) {
            const int p = it.pt, ty = d->pt_type[p];
            const bool series = (ty == AF_BC_QVAR || ty == AF_BC_HVAR);
            J.push_back(it.j);
            I.push_back(it.i);
            VAL.push_back(ty == AF_BC_FREE ? 0.0 : d->pt_value[p]);
            OFF.push_back(series && d->pt_series_off ? d->pt_series_off[p] : 0);
            LEN.push_back(series && d->pt_series_len ? d->pt_series_len[p] : 0);
        }
    };
    build(src, h.src_begin, h.src_j, h.src_i, h.src_val, h.src_off, h.src_len);
    build(ovr, h.ovr_begin, h.ovr_j, h.ovr_i, h.ovr_val, h.ovr_off, h.ovr_len);
    h.n_series = d->n_series;
    h.n_rain = d->rain_enabled ? d->n_rain : 0;
}

int configure(af_domain* D, int precision, const af_tuning* tun) {
    AF_REQUIRE(precision == AF_FP32 || precision == AF_FP64, AF_E_INVALID, "precision %d",
               precision);
    D->precision = precision;
    D->V = precision == AF_FP64 ? 2 : 4;
    D->W = 64 * D->V;
    // Defaults measured on the 16384 x 16384 roofline case: 64-row tiles with
    // two-row bursts at FP32, 32-row tiles at FP64, where the variant follows
    // the wet area (adapt_variant).
    D->R = precision == AF_FP64 ? 32 : 64;
    D->variant = precision == AF_FP64 ? 0 : 1;
    D->variant_auto = (precision == AF_FP64);
    D->wpb = 4;
    D->graph_steps = 36;
    D->dt_log_cap = 0;
    D->report_cap = 256;
    if (tun) {
        if (tun->tile_rows > 0) D->R = D->R_fixed = tun->tile_rows;
        if (tun->waves_per_block > 0) D->wpb = tun->waves_per_block;
        if (tun->graph_steps > 0) D->graph_steps = tun->graph_steps;
        if (tun->dt_log_cap > 0) D->dt_log_cap = tun->dt_log_cap;
        if (tun->report_cap > 0) D->report_cap = tun->report_cap;
        AF_REQUIRE(tun->variant >= 0 && tun->variant <= 3, AF_E_INVALID,
                   "variant %d (0 = default, 1..3)", tun->variant);
        if (tun->variant > 0) D->variant = tun->variant - 1;
        D->variant_auto = (precision == AF_FP64 && tun->variant == 0);
        AF_REQUIRE(tun->tile_order == 0 || tun->tile_order == 1, AF_E_INVALID,
                   "tile_order %d (0 or 1)", tun->tile_order);
        D->tile_order = tun->tile_order;
    }
    AF_REQUIRE(D->R >= 1 && D->R <= 65536, AF_E_INVALID, "tile_rows %d", D->R);
    AF_REQUIRE(D->wpb == 1 || D->wpb == 4, AF_E_INVALID, "waves_per_block %d (1 or 4)", D->wpb);
    AF_REQUIRE(D->graph_steps % 6 == 0 && D->graph_steps <= 6000, AF_E_INVALID,
               "graph_steps %d must be a positive multiple of 6", D->graph_steps);
    return AF_OK;
}

// Table region at the start of the arena: scenario table, tile table, the
// volume-reduction scratch, and the control-block gather list and buffer.
size_t plan_tables(af_domain* D, int max_scn, int max_tiles) {
    Layout L;
    D->o_scns = L.take(sizeof(ScnDev) * (size_t)max_scn);
    D->o_tiles = L.take(sizeof(TileDesc) * (size_t)max_tiles);
    D->o_lane = L.take(sizeof(TileDesc) * (size_t)max_tiles);
    D->o_volbuf = L.take(sizeof(double) * 3 * kVolBlocks);
    D->o_ctlptr = L.take(sizeof(Ctl*) * (size_t)max_scn);
    D->o_ctlbuf = L.take(sizeof(Ctl) * (size_t)max_scn);
    return L.top;
}

// Device bytes of one scenario (its heap extent); fills h's geometry and
// extent-relative offsets.
size_t scenario_bytes(const af_domain* D, const af_scenario_desc* d, HostScn& h) {
    plan_geometry(h, d, D);
    Layout L;
    plan_layout(h, d, D, L);
    return align_up(L.top, 256);
}

bool heap_alloc(af_domain* D, size_t bytes, size_t* off) {
    for (size_t i = 0; i < D->free_list.size(); i++) {
        auto& f = D->free_list[i];
        if (f.second >= bytes) {
            *off = f.first;
            f.first += bytes;
            f.second -= bytes;
            if (f.second == 0) D->free_list.erase(D->free_list.begin() + i);
            return true;
        }
    }
    return false;
}

void heap_free(af_domain* D, size_t off, size_t bytes) {
    auto& fl = D->free_list;
    auto it = std::lower_bound(fl.begin(), fl.end(), std::make_pair(off, (size_t)0));
    it = fl.insert(it, {off, bytes});
    size_t i = it - fl.begin();
    if (i + 1 < fl.size() && fl[i].first + fl[i].second == fl[i + 1].first) {
        fl[i].second += fl[i + 1].second;
        fl.erase(fl.begin() + i + 1);
    }
    if (i > 0 && fl[i - 1].first + fl[i - 1].second == fl[i].first) {
        fl[i - 1].second += fl[i].second;
        fl.erase(fl.begin() + i);
    }
}

size_t heap_free_bytes(const af_domain* D) {
    size_t n = 0;
    for (auto& f : D->free_list) n += f.second;
    return n;
}

// The FP32 path keeps the terrain as 32-bit fixed point: integer units of
// the smallest power-of-two step 2^k (metres) that spans the valid cells in
// at most kTerrainUnits units, above a base that is a multiple of the step.
// Elevations on any grid of 2^-m metres (m <= -k) are exact; others round to
// half a step (3e-8 m for 80 m of relief, where a float elevation would round
// to 4e-6 m).
constexpr double kTerrainUnits = 2147483000.0;   // below INT_MAX; INT_MIN is NODATA

void terrain_scale(const af_scenario_desc* d, double& zbase, double& zscale) {
    bool any = false;
    double lo = 0.0, hi = 0.0;
    for (long long c = 0; c < (long long)d->nx * d->ny; c++) {
        const double v = d->dem[c];
        if (v == d->nodata || !std::isfinite(v)) continue;
        if (!any || v < lo) lo = v;
        if (!any || v > hi) hi = v;
        any = true;
    }
    int k = -40;   // a flat or empty terrain keeps a fine step
    if (any && hi > lo) k = std::max(k, (int)std::ceil(std::log2((hi - lo) / kTerrainUnits)));
    for (;; k++) {
        zscale = std::ldexp(1.0, k);
        zbase = any ? std::floor(lo / zscale) * zscale : 0.0;
        if (!any || (hi - zbase) / zscale <= kTerrainUnits) break;
    }
}

// The DEM and the initial depths in the device layout and precision (the
// terrain as Z: double elevations at FP64, fixed point at FP32).  Returns
// AF_E_INVALID for a non-finite DEM value (other than a non-finite NODATA
// value itself) or initial depth: such a cell would turn its neighbours'
// fluxes to NaN on the first step.
template <typename T, typename Z>
int convert_inputs(const af_scenario_desc* d, long long P, Z* z, T* H, double& zbase,
                   double& zscale, double& hmax, long long& wet) {
    constexpr bool f64 = sizeof(T) == 8;
    zbase = 0.0;
    zscale = 1.0;
    if (!f64) terrain_scale(d, zbase, zscale);
    memset(z, 0, sizeof(Z) * (size_t)P * d->ny);
    memset(H, 0, sizeof(T) * (size_t)P * d->ny);
    hmax = -DBL_MAX;
    wet = 0;
    for (int j = 0; j < d->ny; j++) {
        for (int i = 0; i < d->nx; i++) {
            const long long c = i + (long long)j * d->nx;
            const double v = d->dem[c];
            const bool nod = (v == d->nodata);
            AF_REQUIRE(nod || std::isfinite(v), AF_E_INVALID,
                       "DEM value at column %d, row %d is not finite (%g)", i, j, v);
            Z zt;
            if constexpr (f64) zt = (Z)v;
            else zt = nod ? (Z)INT_MIN : (Z)std::llround((v - zbase) / zscale);
            z[(size_t)j * P + i] = zt;
            T ht = d->h0 ? (T)d->h0[c] : T(0);
            AF_REQUIRE(std::isfinite((double)ht), AF_E_INVALID,
                       "initial depth at column %d, row %d is not finite (%g)", i, j,
                       d->h0[c]);
            H[(size_t)j * P + i] = ht;
            if (!nod && (double)ht > hmax) hmax = (double)ht;  // kernels.cpp:32
            if ((double)ht > 0.01) wet++;
        }
    }
    return AF_OK;
}

int stage_inputs(const af_scenario_desc* d, int precision, Staged& s) {
    int rc = validate(d);
    if (rc) return rc;
    AF_REQUIRE(precision == AF_FP32 || precision == AF_FP64, AF_E_INVALID, "precision %d",
               precision);
    s.precision = precision;
    s.nx = d->nx;
    s.ny = d->ny;
    s.P = row_pitch(d->nx, precision);
    const size_t bytes = (size_t)s.P * d->ny * elem_size(precision);
    AF_REQUIRE(s.z.reserve(bytes) && s.h.reserve(bytes), AF_E_CAPACITY,
               "stage: no host memory for %zu bytes", 2 * bytes);
    if (precision == AF_FP64)
        return convert_inputs<double, double>(d, s.P, (double*)s.z.data(), (double*)s.h.data(),
                                              s.zbase, s.zscale, s.hmax, s.wet);
    return convert_inputs<float, int>(d, s.P, (int*)s.z.data(), (float*)s.h.data(), s.zbase,
                                      s.zscale, s.hmax, s.wet);
}

// Zeroes `words` 4-byte words at p (16-byte aligned): 16-byte stores for the
// body, single words for the tail.  (hipMemsetAsync of a few MiB measured
// about 1 ms to enqueue on the host, longer than the fill itself.)
__global__ void __launch_bounds__(256) k_zero(uint4* __restrict__ p, long long words) {
    const long long n4 = words >> 2;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
    const long long t = (n4 << 2) + blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < words) reinterpret_cast<unsigned*>(p)[t] = 0u;
}

int zero_fill(void* p, size_t bytes, hipStream_t st) {
    AF_REQUIRE(((uintptr_t)p & 15) == 0 && bytes % 4 == 0, AF_E_INVALID,
               "zero_fill: unaligned range");
    const long long words = (long long)(bytes / 4);
    if (words == 0) return AF_OK;
    const long long blocks = std::min<long long>((words / 4 + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st,
                       reinterpret_cast<uint4*>(p), words);
    AF_HIP(hipGetLastError());
    return AF_OK;
}

int upload_scenario(af_domain* D, int s, const af_scenario_desc* d, hipStream_t st,
                    int parity, long long kmod, const Staged& pre) {
    HostScn& h = D->scn[s];
    AF_REQUIRE(pre.precision == D->precision && pre.nx == d->nx && pre.ny == d->ny &&
                   pre.P == h.P,
               AF_E_INVALID, "staged scenario (%d x %d, %s) does not match the pool (%s)",
               pre.nx, pre.ny, pre.precision == AF_FP64 ? "fp64" : "fp32",
               D->precision == AF_FP64 ? "fp64" : "fp32");
    plan_points(h, d, D);
    char* A = D->arena;
    // device pointers
    ScnDev& v = h.dev;
    v = ScnDev{};
    v.nx = h.nx;
    v.ny = h.ny;
    v.P = h.P;
    v.nstrips = h.nstrips;
    v.tile_rows = h.R;
    v.tile_base = 0;
    v.ntiles = h.ntiles;
    v.z = A + h.o_z;
    for (int b = 0; b < 2; b++) {
        v.H[b] = A + h.o_H[b];
        v.Qx[b] = A + h.o_Qx[b];
        v.Qy[b] = A + h.o_Qy[b];
    }
    v.dx = d->dx;
    v.dxinv = 1.0 / d->dx;
    v.dA = d->dx * d->dx;
    v.fn = d->fpn;
    v.thresh = d->depth_thresh;
    v.maxh = d->max_hflow;
    v.init_tstep = d->init_tstep;
    v.cfl = d->cfl;
    v.g = d->g;
    v.theta = d->theta;
    v.sim_time = d->sim_time;
    v.mass_int = d->mass_int;
    v.nodata_t = d->nodata;   // the FP32 terrain marks NODATA with INT_MIN instead
    v.west_free = d->west_free ? 1 : 0;
    v.east_free = d->east_free ? 1 : 0;
    v.north_free = d->north_free ? 1 : 0;
    v.south_free = d->south_free ? 1 : 0;
    v.src_begin = reinterpret_cast<const int*>(A + h.o_src_begin);
    v.src_j = reinterpret_cast<const int*>(A + h.o_src_j);
    v.src_i = reinterpret_cast<const int*>(A + h.o_src_i);
    v.src_val = reinterpret_cast<const double*>(A + h.o_src_val);
    v.src_off = reinterpret_cast<const int*>(A + h.o_src_off);
    v.src_len = reinterpret_cast<const int*>(A + h.o_src_len);
    v.ovr_begin = reinterpret_cast<const int*>(A + h.o_ovr_begin);
    v.ovr_j = reinterpret_cast<const int*>(A + h.o_ovr_j);
    v.ovr_i = reinterpret_cast<const int*>(A + h.o_ovr_i);
    v.ovr_val = reinterpret_cast<const double*>(A + h.o_ovr_val);
    v.ovr_off = reinterpret_cast<const int*>(A + h.o_ovr_off);
    v.ovr_len = reinterpret_cast<const int*>(A + h.o_ovr_len);
    v.series_t = reinterpret_cast<const double*>(A + h.o_series_t);
    v.series_v = reinterpret_cast<const double*>(A + h.o_series_v);
    v.rain_enabled = d->rain_enabled ? 1 : 0;
    v.n_rain = h.n_rain;
    v.rain_t = reinterpret_cast<const double*>(A + h.o_rain_t);
    v.rain_r = reinterpret_cast<const double*>(A + h.o_rain_r);
    v.ctl = reinterpret_cast<Ctl*>(A + h.o_ctl);
    v.part_vol = reinterpret_cast<double*>(A + h.o_part_vol);
    v.part_comp = reinterpret_cast<double*>(A + h.o_part_comp);
    v.part_area = reinterpret_cast<unsigned long long*>(A + h.o_part_area);
    v.dt_log = reinterpret_cast<double*>(A + h.o_dt_log);
    v.dt_log_cap = D->dt_log_cap;
    v.reports = reinterpret_cast<double*>(A + h.o_reports);
    v.report_cap = D->report_cap;

    // state: the terrain and the initial depths (buffer `parity`) from the
    // staged inputs, every other state buffer zero (one fill over the
    // contiguous H, Qx, Qy extents of both buffers, before the depth copy)
    const size_t es = elem_size(D->precision);
    const size_t cells = (size_t)h.P * h.ny, faces_y = (size_t)h.P * (h.ny + 1);
    AF_REQUIRE(h.o_H[0] < h.o_Qx[0] && h.o_Qy[0] < h.o_H[1] && h.o_Qx[1] < h.o_Qy[1],
               AF_E_INVALID, "state layout out of order");
    int zrc = zero_fill(A + h.o_H[0], h.o_Qy[1] + faces_y * es - h.o_H[0], st);
    if (zrc) return zrc;
    const double hmax = pre.hmax;
    h.wet0 = pre.wet;
    h.zbase = pre.zbase;
    h.zscale = pre.zscale;
    AF_HIP(hipMemcpyAsync(A + h.o_z, pre.z.data(), cells * es, hipMemcpyHostToDevice, st));
    AF_HIP(hipMemcpyAsync(A + h.o_H[parity], pre.h.data(), cells * es, hipMemcpyHostToDevice, st));
    v.zscale = h.zscale;
    v.zbase = h.zbase;

    // point lists, series, rain and the control block: consecutive extents of
    // the layout (src_begin first, the control block last), assembled on the
    // host and uploaded with one copy
    const size_t aux_lo = h.o_src_begin, aux_hi = h.o_ctl + sizeof(Ctl);
    std::vector<char> aux(aux_hi - aux_lo, 0);
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) memcpy(aux.data() + (off - aux_lo), src, bytes);
    };
    put(h.o_src_begin, h.src_begin.data(), h.src_begin.size() * 4);
    put(h.o_src_j, h.src_j.data(), h.src_j.size() * 4);
    put(h.o_src_i, h.src_i.data(), h.src_i.size() * 4);
    put(h.o_src_val, h.src_val.data(), h.src_val.size() * 8);
    put(h.o_src_off, h.src_off.data(), h.src_off.size() * 4);
    put(h.o_src_len, h.src_len.data(), h.src_len.size() * 4);
    put(h.o_ovr_begin, h.ovr_begin.data(), h.ovr_begin.size() * 4);
    put(h.o_ovr_j, h.ovr_j.data(), h.ovr_j.size() * 4);
    put(h.o_ovr_i, h.ovr_i.data(), h.ovr_i.size() * 4);
    put(h.o_ovr_val, h.ovr_val.data(), h.ovr_val.size() * 8);
    put(h.o_ovr_off, h.ovr_off.data(), h.ovr_off.size() * 4);
    put(h.o_ovr_len, h.ovr_len.data(), h.ovr_len.size() * 4);
    put(h.o_series_t, d->series_t, (size_t)h.n_series * 8);
    put(h.o_series_v, d->series_v, (size_t)h.n_series * 8);
    put(h.o_rain_t, d->rain_t, (size_t)h.n_rain * 8);
    put(h.o_rain_r, d->rain_r, (size_t)h.n_rain * 8);

    Ctl c{};
    memset(&c, 0, sizeof(c));
    const int cur = parity;
    c.t[cur] = 0.0;                       // main.cpp:197
    c.next_mass[cur] = d->mass_int;       // main.cpp:200
    const int hs = (int)(kmod % 3);
    c.hmax[hs] = host_key(hmax);
    c.hmax[(hs + 1) % 3] = host_key(-DBL_MAX);
    c.hmax[(hs + 2) % 3] = host_key(-DBL_MAX);
    c.t[cur ^ 1] = 0.0;
    c.next_mass[cur ^ 1] = d->mass_int;
    c.par[0] = c.par[1] = parity;  // the initial state sits in buffer `parity`
    c.last_wet = (unsigned long long)h.wet0;
    put(h.o_ctl, &c, sizeof(c));
    AF_HIP(hipMemcpyAsync(A + aux_lo, aux.data(), aux.size(), hipMemcpyHostToDevice, st));
    // the host sources (aux, the staged inputs) must outlive the copies
    AF_HIP(hipStreamSynchronize(st));
    return AF_OK;
}

// Dynamic LDS a k_step workgroup reserves so that at most OCC waves per SIMD
// (OCC * 4 / WPB workgroups per CU) run at once.  The register allocation
// alone can admit more: the FP64 step compiles to 168 VGPRs, which lets three
// waves per SIMD in, and three measured 12% slower than two on the 16384^2
// step (3.49 against 3.11 ms per launch, same box).  Reserving just over
// 1/(n+1) of the CU's LDS per workgroup leaves room for n of them and not for
// n+1.  Returns 0 where the reservation would exceed a workgroup's limit.
template <typename T, int WPB, int OCC>
unsigned occupancy_lds(const af_domain* D) {
    const size_t per_cu = (size_t)OCC * 4 / WPB;
    const size_t need = (size_t)D->lds_cu / (per_cu + 1) + 1;
    const size_t have = step_static_lds<T, WPB>();
    if (D->lds_cu <= 0 || need <= have || need > (size_t)D->lds_block) return 0;
    return (unsigned)(need - have);
}

template <typename T, int V, bool F64, int OCC, int PF>
int launch_step_v(af_domain* D, const TileDesc* tiles, int ntiles, int kmod6, hipStream_t st) {
    const int wpb = D->wpb;
    const unsigned blocks = (unsigned)((ntiles + wpb - 1) / wpb);
    if (wpb == 4) {
        const unsigned lds = occupancy_lds<T, 4, OCC>(D);
        hipLaunchKernelGGL((k_step<T, V, F64, 4, OCC, PF>), dim3(blocks), dim3(256), lds, st,
                           D->d_scns, tiles, ntiles, kmod6);
    } else {
        const unsigned lds = occupancy_lds<T, 1, OCC>(D);
        hipLaunchKernelGGL((k_step<T, V, F64, 1, OCC, PF>), dim3(blocks), dim3(64), lds, st,
                           D->d_scns, tiles, ntiles, kmod6);
    }
    AF_HIP(hipGetLastError());
    return AF_OK;
}

// Kernel variants (af_tuning.variant - 1): rows per loop step (PF = 1: one
// row, the next in flight; PF = 2: two-row load and store bursts) and the
// occupancy (waves per SIMD) the register allocation targets and the launch
// holds (occupancy_lds).
int launch_tiles(af_domain* D, const TileDesc* t, int n, int kmod6, hipStream_t st) {
    if (n == 0) return AF_OK;
    if (D->precision == AF_FP64) {
        switch (D->variant) {
            case 1: return launch_step_v<double, 2, true, 2, 2>(D, t, n, kmod6, st);
            case 2: return launch_step_v<double, 2, true, 1, 2>(D, t, n, kmod6, st);
            default: return launch_step_v<double, 2, true, 3, 1>(D, t, n, kmod6, st);
        }
    }
    switch (D->variant) {
        case 1: return launch_step_v<float, 4, false, 3, 2>(D, t, n, kmod6, st);
        case 2: return launch_step_v<float, 4, false, 2, 2>(D, t, n, kmod6, st);
        default: return launch_step_v<float, 4, false, 4, 1>(D, t, n, kmod6, st);
    }
}

// One step of every active scenario.
int launch_step(af_domain* D, int kmod6, hipStream_t st) {
    return launch_tiles(D, D->d_tiles, D->ntiles, kmod6, st);
}

// Captures `steps` launches over the tile table (tiles, n), with step phases
// 0, 1, 2, ... (mod 6); `steps` is a multiple of 6, so a replay leaves every
// scenario's rotating slots where it found them.
int capture_steps(af_domain* D, const TileDesc* tiles, int n, int steps, hipGraph_t* gout,
                  hipGraphExec_t* eout) {
    hipStream_t cs;
    AF_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
        return AF_E_HIP;
    }
    int rc = AF_OK;
    for (int i = 0; i < steps && rc == AF_OK; i++) rc = launch_tiles(D, tiles, n, i % 6, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != AF_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != AF_OK) return rc;
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return AF_E_HIP;
    }
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        return AF_E_HIP;
    }
    *gout = g;
    *eout = ge;
    return AF_OK;
}

void drop_graph(hipGraph_t& g, hipGraphExec_t& e) {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
    e = nullptr;
    g = nullptr;
}

int build_graph(af_domain* D, hipStream_t st) {
    (void)st;
    if (D->gexec) return AF_OK;
    return capture_steps(D, D->d_tiles, D->ntiles, D->graph_steps, &D->graph, &D->gexec);
}

// Reduction of sum(H*dA) and wet area for af_domain_volume.
template <typename T>
__global__ void __launch_bounds__(256)
k_volume(const T* __restrict__ H, long long P, int nx, int ny, double dA, double* __restrict__ out) {
    double s = 0.0, c = 0.0;
    double wet = 0.0;
    const long long n = (long long)nx * ny;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        const long long j = t / nx, i = t % nx;
        const double h = (double)H[j * P + i];
        kahan_add(s, c, h * dA);
        if (h > 0.01) wet += 1.0;
    }
    __shared__ double ss[256], sc2[256], sw[256];
    ss[threadIdx.x] = s;
    sc2[threadIdx.x] = c;
    sw[threadIdx.x] = wet;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, w = 0.0;
        for (int k = 0; k < 256; k++) {
            kahan_add(a, b, ss[k]);
            b += sc2[k];
            w += sw[k];
        }
        out[3 * blockIdx.x + 0] = a;
        out[3 * blockIdx.x + 1] = b;
        out[3 * blockIdx.x + 2] = w;
    }
}

int current_parity(af_domain* D, int s, long long* n_done, hipStream_t st) {
    Ctl c;
    AF_HIP(hipMemcpyAsync(&c, D->scn[s].dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    if (n_done) *n_done = c.n_done;
    return (int)((D->scn[s].base_parity + c.n_done) & 1);
}

__global__ void k_finalize_reports(const ScnDev* __restrict__ scn) {
    Stepper<double, 2, true>::finalize_report(*scn, threadIdx.x & 63);
}

// Completes a mass report whose partial sums the last executed step left.
int finalize_reports(af_domain* D, int s, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize_reports, dim3(1), dim3(64), 0, st, D->d_scns + s);
    AF_HIP(hipGetLastError());
    return AF_OK;
}

static_assert(sizeof(Ctl) % 8 == 0, "Ctl is copied in 8-byte words");
constexpr int kCtlWords = (int)(sizeof(Ctl) / 8);

// Copies the control blocks of the n listed scenarios into one contiguous
// buffer, so that the host reads all of them with one copy.
__global__ void k_gather_ctl(Ctl* const* __restrict__ list, int n, Ctl* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kCtlWords) return;
    const int s = i / kCtlWords, w = i - s * kCtlWords;
    reinterpret_cast<unsigned long long*>(out)[i] =
        reinterpret_cast<const unsigned long long*>(list[s])[w];
}

// The control blocks of the active scenarios (slots in D->active order),
// read with one gather launch, one copy and one synchronisation.
int settle(af_domain* D, hipStream_t st);

// The FP64 default variant follows the flow.  The exact pow(h, 10/3) runs in
// every wavefront with a wet face, and more waves per SIMD hide its chain:
// the one-row loop at three waves (variant 1) is the faster on a wet flow and
// on small and mid-size domains; the two-row bursts at two waves (variant 2)
// only on a large, nearly dry one.  Measured on the 16384^2 roofline case
// (ms per step): over its first 110 steps 3.10 (variant 2) against 3.30;
// over the full 3600 s run 3.55 against 3.20, variant 1 ahead from between
// t = 300 and 500 s, when 2.5 to 3% of the cells are wet.  Variant 1 is also the
// faster on the 4096^2 valley (0.288 against 0.320), the ragged ensemble
// and the served pool; variant 2 on a dry 8192^2 roofline case (0.813
// against 0.868).  So variant 2 is used while the largest scenario has at
// least kBurstTiles tiles and under kBurstWet of all cells are wet (a pool
// uses variant 1).  The wet count is the last mass report's (or the initial
// state's); the choice is made wherever the control blocks are read back,
// and a change drops the captured graphs, which the next advance captures
// again.  The variants compute bitwise the same results.
constexpr int kBurstTiles = 8192;
constexpr double kBurstWet = 0.03;

void adapt_variant(af_domain* D, double wet, double cells) {
    if (!D->variant_auto) return;
    const int want = (D->most_tiles >= kBurstTiles && wet < kBurstWet * cells) ? 1 : 0;
    if (want == D->variant) return;
    D->variant = want;
    drop_graph(D->graph, D->gexec);
    drop_graph(D->lgraph, D->lgexec);
}

int gather_ctl(af_domain* D, std::vector<Ctl>& out, hipStream_t st) {
    int rc = settle(D, st);
    if (rc) return rc;
    const int n = (int)D->active.size();
    out.resize(n);
    if (n == 0) return AF_OK;
    const int words = n * kCtlWords;
    hipLaunchKernelGGL(k_gather_ctl, dim3((words + 255) / 256), dim3(256), 0, st, D->d_ctlptr, n,
                       D->d_ctlbuf);
    AF_HIP(hipGetLastError());
    AF_HIP(hipMemcpyAsync(out.data(), D->d_ctlbuf, sizeof(Ctl) * n, hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    double wet = 0.0, cells = 0.0;
    for (int a = 0; a < n; a++) {
        wet += (double)out[a].last_wet;
        cells += (double)D->scn[D->active[a]].nx * D->scn[D->active[a]].ny;
    }
    adapt_variant(D, wet, cells);
    return AF_OK;
}

// A slot is finished when its clock reached sim_time (main.cpp:227) or a
// non-zero status stopped it.
bool slot_finished(const af_domain* D, int s, const Ctl& c, int par) {
    return !(c.t[par] < D->scn[s].dev.sim_time) || c.status[par] != 0;
}

// The step index whose phase (mod 6) slot s's next step takes: the pool's,
// or 0 for a scenario in the lane, which entered the lane at a pool step
// index that is a multiple of 6 and steps there in multiples of 6.
long long scn_k(const af_domain* D, int s) { return D->scn[s].in_lane ? 0 : D->k; }

// Makes `st` wait for the lane work enqueued so far: every call that reads
// or changes scenario state or the tile tables does this first.
hipStream_t joined(af_domain* D, af_stream_t stream) {
    hipStream_t st = to_stream(stream);
    if (D->lane_pending) {
        if (hipStreamWaitEvent(st, D->ev_join, 0) != hipSuccess)
            (void)hipStreamSynchronize(D->lane_st);
        D->lane_pending = false;
    }
    return st;
}

const char* status_text(int status) {
    if (status & kStatusNonfinite) return "a step produced a NaN or infinite depth";
    if (status & kStatusStalled) return "the clock stalled (t + dt == t)";
    return "unknown status";
}

// The lane's tile table: every tile of the lane's active scenarios, listed as
// the full table lists them.  Released slots leave the lane.  The lane's
// captured graph holds the tile count, so it is dropped.
int build_lane(af_domain* D, hipStream_t st) {
    std::vector<int> keep;
    for (int s : D->lane_slots)
        if (s >= 0 && s < D->max_scn && D->scn[s].active && D->scn[s].in_lane) keep.push_back(s);
    D->lane_slots.swap(keep);
    std::vector<TileDesc> tiles;
    for (int s : D->lane_slots) {
        const HostScn& h = D->scn[s];
        for (int seg = 0; seg < h.nsegs; seg++)
            for (int strip = 0; strip < h.nstrips; strip++)
                tiles.push_back({s, strip, seg, (strip == 0 && seg == 0) ? 1 : 0});
    }
    if (!tiles.empty())
        AF_HIP(hipMemcpyAsync(D->d_lane, tiles.data(), sizeof(TileDesc) * tiles.size(),
                              hipMemcpyHostToDevice, st));
    AF_HIP(hipStreamSynchronize(st));
    D->nlane = (int)tiles.size();
    drop_graph(D->lgraph, D->lgexec);
    return AF_OK;
}

// Rebuilds and uploads the tile table from the active scenarios; a captured
// graph holds the tile count, so it is dropped.
int rebuild_tiles(af_domain* D, hipStream_t st) {
    std::vector<TileDesc> tiles;
    std::vector<Ctl*> ctls;
    std::vector<int> act;
    for (int s = 0; s < D->max_scn; s++) {
        const HostScn& h = D->scn[s];
        if (!h.active) continue;
        act.push_back(s);
        ctls.push_back(h.dev.ctl);
        if (h.in_lane) continue;    // stepped by the lane
        for (int seg = 0; seg < h.nsegs; seg++)
            for (int strip = 0; strip < h.nstrips; strip++)
                tiles.push_back({s, strip, seg, (strip == 0 && seg == 0) ? 1 : 0});
    }
    AF_REQUIRE((int)tiles.size() <= D->max_tiles, AF_E_CAPACITY, "tile table full (%zu > %d)",
               tiles.size(), D->max_tiles);
    if (D->tile_order == 1) {
        // Largest first: workgroups are dispatched roughly in table order as
        // slots free up, so the step's tail is made of the short tiles (the
        // last segment of a scenario, the partial strip of a ragged width).
        // The kernel indexes per-scenario data by (strip, segment), so the
        // order of the table does not change any result.
        auto cells = [&](const TileDesc& t) {
            const HostScn& h = D->scn[t.scn];
            const long long rows = std::min(h.R, h.ny - t.seg * h.R);
            const long long cols = std::min(D->W, h.nx - t.strip * D->W);
            return rows * cols;
        };
        std::stable_sort(tiles.begin(), tiles.end(), [&](const TileDesc& a, const TileDesc& b) {
            return cells(a) > cells(b);
        });
    }
    if (!tiles.empty())
        AF_HIP(hipMemcpyAsync(D->d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size(),
                              hipMemcpyHostToDevice, st));
    if (!ctls.empty())
        AF_HIP(hipMemcpyAsync(D->d_ctlptr, ctls.data(), sizeof(Ctl*) * ctls.size(),
                              hipMemcpyHostToDevice, st));
    AF_HIP(hipStreamSynchronize(st));
    D->ntiles = (int)tiles.size();
    D->active = std::move(act);
    D->dirty = false;
    drop_graph(D->graph, D->gexec);
    return build_lane(D, st);
}

// Admissions and releases mark the tile tables stale; the next call that
// launches steps or gathers the control blocks rebuilds them once.
// (Lane work still in flight reads the lane's table: the stream waits for it
// first.)
int settle(af_domain* D, hipStream_t st) {
    if (!D->dirty) return AF_OK;
    if (D->lane_pending) {
        AF_HIP(hipStreamWaitEvent(st, D->ev_join, 0));
        D->lane_pending = false;
    }
    return rebuild_tiles(D, st);
}

// Re-sorts one point list (sources or overrides) of scenario h for its
// present tile height: by (tile, row, column), stably, so entries of one cell
// keep their file order.
void resort_points(const HostScn& h, int W, std::vector<int>& begin, std::vector<int>& J,
                   std::vector<int>& I, std::vector<double>& VAL, std::vector<int>& OFF,
                   std::vector<int>& LEN) {
    const size_t n = J.size();
    std::vector<int> tile(n), idx(n);
    for (size_t e = 0; e < n; e++) {
        tile[e] = (I[e] / W) + (J[e] / h.R) * h.nstrips;
        idx[e] = (int)e;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (tile[a] != tile[b]) return tile[a] < tile[b];
        if (J[a] != J[b]) return J[a] < J[b];
        return I[a] < I[b];
    });
    auto permute = [&](auto& v) {
        auto w = v;
        for (size_t e = 0; e < n; e++) w[e] = v[idx[e]];
        v.swap(w);
    };
    permute(J);
    permute(I);
    permute(VAL);
    permute(OFF);
    permute(LEN);
    begin.assign(h.ntiles + 1, 0);
    for (size_t e = 0; e < n; e++) begin[tile[idx[e]] + 1]++;
    for (int t = 0; t < h.ntiles; t++) begin[t + 1] += begin[t];
}

// Moves active scenario s to tile height R (R_alone <= R): completes a mass
// report pending from the last step (its partials are laid out by the old
// tiles), re-sorts and uploads the point lists, and uploads the scenario's
// constants.  Results do not depend on the tiling.  The caller rebuilds the
// tile table.
int retile(af_domain* D, int s, int R, hipStream_t st) {
    HostScn& h = D->scn[s];
    hipLaunchKernelGGL(k_finalize_reports, dim3(1), dim3(64), 0, st, D->d_scns + s);
    AF_HIP(hipGetLastError());
    h.R = R;
    h.nsegs = (h.ny + R - 1) / R;
    h.ntiles = h.nstrips * h.nsegs;
    resort_points(h, D->W, h.src_begin, h.src_j, h.src_i, h.src_val, h.src_off, h.src_len);
    resort_points(h, D->W, h.ovr_begin, h.ovr_j, h.ovr_i, h.ovr_val, h.ovr_off, h.ovr_len);
    char* A = D->arena;
    auto up = [&](size_t off, const void* src, size_t bytes) -> int {
        if (bytes == 0) return AF_OK;
        AF_HIP(hipMemcpyAsync(A + off, src, bytes, hipMemcpyHostToDevice, st));
        return AF_OK;
    };
    int rc = AF_OK;
    if ((rc = up(h.o_src_begin, h.src_begin.data(), h.src_begin.size() * 4))) return rc;
    if ((rc = up(h.o_src_j, h.src_j.data(), h.src_j.size() * 4))) return rc;
    if ((rc = up(h.o_src_i, h.src_i.data(), h.src_i.size() * 4))) return rc;
    if ((rc = up(h.o_src_val, h.src_val.data(), h.src_val.size() * 8))) return rc;
    if ((rc = up(h.o_src_off, h.src_off.data(), h.src_off.size() * 4))) return rc;
    if ((rc = up(h.o_src_len, h.src_len.data(), h.src_len.size() * 4))) return rc;
    if ((rc = up(h.o_ovr_begin, h.ovr_begin.data(), h.ovr_begin.size() * 4))) return rc;
    if ((rc = up(h.o_ovr_j, h.ovr_j.data(), h.ovr_j.size() * 4))) return rc;
    if ((rc = up(h.o_ovr_i, h.ovr_i.data(), h.ovr_i.size() * 4))) return rc;
    if ((rc = up(h.o_ovr_val, h.ovr_val.data(), h.ovr_val.size() * 8))) return rc;
    if ((rc = up(h.o_ovr_off, h.ovr_off.data(), h.ovr_off.size() * 4))) return rc;
    if ((rc = up(h.o_ovr_len, h.ovr_len.data(), h.ovr_len.size() * 4))) return rc;
    h.dev.tile_rows = R;
    h.dev.ntiles = h.ntiles;
    AF_HIP(hipMemcpyAsync(D->d_scns + s, &h.dev, sizeof(ScnDev), hipMemcpyHostToDevice, st));
    AF_HIP(hipStreamSynchronize(st));  // the host vectors above are sources of the copies
    return AF_OK;
}

// A pool that has drained below kMinTiles tiles runs its scenarios nearly
// alone: those admitted under load, with taller tiles, move to the height
// they would take alone (as long as the tile table has room).
int retile_on_drain(af_domain* D, hipStream_t st) {
    int used = 0;
    for (const HostScn& h : D->scn) used += h.active ? h.ntiles : 0;
    if (used >= kMinTiles) return AF_OK;
    for (int s = 0; s < D->max_scn; s++) {
        HostScn& h = D->scn[s];
        if (!h.active || h.R <= h.R_alone) continue;
        const int grow = h.cap_tiles - h.ntiles;
        if (used + grow > D->max_tiles) continue;
        int rc = retile(D, s, h.R_alone, st);
        if (rc) return rc;
        used += grow;
    }
    return AF_OK;
}

int pool_init(af_domain* D, int precision, const af_tuning* tuning, int max_scn, int max_tiles,
              void* arena, size_t arena_bytes) {
    int rc = configure(D, precision, tuning);
    if (rc) return rc;
    AF_REQUIRE(max_scn >= 1 && max_tiles >= 1, AF_E_INVALID, "pool needs >= 1 scenario and tile");
    AF_REQUIRE(arena && ((uintptr_t)arena & 255) == 0, AF_E_INVALID,
               "arena must be 256-byte aligned device memory");
    const size_t tables = plan_tables(D, max_scn, max_tiles);
    AF_REQUIRE(tables <= arena_bytes, AF_E_CAPACITY, "arena too small: tables need %zu bytes",
               tables);
    D->max_scn = max_scn;
    D->max_tiles = max_tiles;
    D->scn.assign(max_scn, HostScn{});
    D->arena = reinterpret_cast<char*>(arena);
    D->arena_bytes = arena_bytes;
    D->heap_lo = tables;
    D->free_list.clear();
    if (arena_bytes > tables) D->free_list.push_back({tables, arena_bytes - tables});
    D->d_scns = reinterpret_cast<ScnDev*>(D->arena + D->o_scns);
    D->d_tiles = reinterpret_cast<TileDesc*>(D->arena + D->o_tiles);
    D->d_lane = reinterpret_cast<TileDesc*>(D->arena + D->o_lane);
    D->d_volbuf = reinterpret_cast<double*>(D->arena + D->o_volbuf);
    D->d_ctlptr = reinterpret_cast<Ctl**>(D->arena + D->o_ctlptr);
    D->d_ctlbuf = reinterpret_cast<Ctl*>(D->arena + D->o_ctlbuf);
    D->active.clear();
    // LDS limits of the device the domain lives on, for the occupancy cap of
    // occupancy_lds; without them the launches reserve nothing
    int dev = 0;
    D->lds_cu = D->lds_block = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&D->lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              dev) != hipSuccess ||
        hipDeviceGetAttribute(&D->lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess) {
        D->lds_cu = D->lds_block = 0;
        (void)hipGetLastError();   // the failed query is not the next launch's error
    }
    return AF_OK;
}

// Places a scenario in a free slot and heap extent and uploads it; its state
// enters at the pool's current step index k.
// Admits scenario d into a free slot.  `pre` holds its inputs staged by
// stage_inputs (which also validated d); without it they are staged here.
int pool_admit(af_domain* D, const af_scenario_desc* d, hipStream_t st, bool rebuild, int* slot,
               const Staged* pre = nullptr) {
    int rc = AF_OK;
    Staged local;
    if (!pre) {
        rc = stage_inputs(d, D->precision, local);
        if (rc) return rc;
        pre = &local;
    }
    int s = -1;
    for (int i = 0; i < D->max_scn; i++)
        if (!D->scn[i].active) { s = i; break; }
    AF_REQUIRE(s >= 0, AF_E_CAPACITY, "no free scenario slot (%d in use)", D->max_scn);
    int used = 0;
    for (const HostScn& o : D->scn) used += o.active ? o.ntiles : 0;
    if (D->pool) D->min_rows = (used < kMinTiles) ? kMinRowsAlone : kMinRowsShared;
    HostScn h;
    const size_t bytes = scenario_bytes(D, d, h);
    AF_REQUIRE(used + h.ntiles <= D->max_tiles, AF_E_CAPACITY,
               "tile table full: %d in use, %d more needed, %d max", used, h.ntiles,
               D->max_tiles);
    size_t off = 0;
    AF_REQUIRE(heap_alloc(D, bytes, &off), AF_E_CAPACITY,
               "arena full: %zu bytes needed, %zu free", bytes, heap_free_bytes(D));
    rebase(h, off);
    h.bytes = bytes;
    D->scn[s] = std::move(h);
    HostScn& hs = D->scn[s];
    const int parity = (int)(D->k & 1);
    rc = upload_scenario(D, s, d, st, parity, D->k % 6, *pre);
    if (rc) {
        heap_free(D, off, bytes);
        D->scn[s] = HostScn{};
        return rc;
    }
    hs.base_parity = parity;
    hs.active = true;
    AF_HIP(hipMemcpyAsync(D->d_scns + s, &hs.dev, sizeof(ScnDev), hipMemcpyHostToDevice, st));
    AF_HIP(hipStreamSynchronize(st));
    if (rebuild) D->dirty = true;
    *slot = s;
    return AF_OK;
}

}  // namespace
}  // namespace af

using namespace af;

extern "C" {

int af_batch_arena_bytes(const af_scenario_desc* descs, int n, int precision,
                         const af_tuning* tuning, size_t* bytes) {
    AF_REQUIRE(descs && n >= 1 && bytes, AF_E_INVALID, "arena_bytes: bad arguments");
    af_domain tmp;
    int rc = configure(&tmp, precision, tuning);
    if (rc) return rc;
    tmp.min_rows = (n == 1) ? kMinRowsAlone : kMinRowsShared;
    size_t heap = 0;
    int tiles = 0;
    for (int s = 0; s < n; s++) {
        rc = validate(&descs[s]);
        if (rc) return rc;
        HostScn h;
        heap += scenario_bytes(&tmp, &descs[s], h);
        tiles += h.ntiles;
    }
    *bytes = plan_tables(&tmp, n, tiles) + heap;
    return AF_OK;
}

int af_domain_arena_bytes(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                          size_t* bytes) {
    return af_batch_arena_bytes(desc, 1, precision, tuning, bytes);
}

int af_batch_create(const af_scenario_desc* descs, int n, int precision, const af_tuning* tuning,
                    void* arena, size_t arena_bytes, af_stream_t stream, af_domain** out) {
    AF_REQUIRE(descs && n >= 1 && arena && out, AF_E_INVALID, "create: bad arguments");
    size_t need = 0;
    int rc = af_batch_arena_bytes(descs, n, precision, tuning, &need);
    if (rc) return rc;
    AF_REQUIRE(need <= arena_bytes, AF_E_CAPACITY, "arena too small: need %zu bytes, got %zu",
               need, arena_bytes);
    af_domain tmp;
    (void)configure(&tmp, precision, tuning);
    tmp.min_rows = (n == 1) ? kMinRowsAlone : kMinRowsShared;
    int tiles = 0, most = 0;
    for (int s = 0; s < n; s++) {
        HostScn h;
        (void)scenario_bytes(&tmp, &descs[s], h);
        tiles += h.ntiles;
        most = std::max(most, h.ntiles);
    }
    af_domain* D = new af_domain();
    hipStream_t st = to_stream(stream);
    rc = pool_init(D, precision, tuning, n, tiles, arena, arena_bytes);
    D->min_rows = tmp.min_rows;
    double wet = 0.0, cells = 0.0;
    for (int s = 0; s < n && rc == AF_OK; s++) {
        int slot = -1;
        rc = pool_admit(D, &descs[s], st, false, &slot);
        if (rc == AF_OK) {
            wet += (double)D->scn[slot].wet0;
            cells += (double)descs[s].nx * descs[s].ny;
        }
    }
    D->most_tiles = most;
    if (rc == AF_OK) adapt_variant(D, wet, cells);
    if (rc == AF_OK) rc = rebuild_tiles(D, st);
    if (rc) {
        delete D;
        return rc;
    }
    *out = D;
    return AF_OK;
}

int af_domain_create(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                     void* arena, size_t arena_bytes, af_stream_t stream, af_domain** out) {
    return af_batch_create(desc, 1, precision, tuning, arena, arena_bytes, stream, out);
}

int af_pool_arena_bytes(int max_scenarios, int max_tiles, size_t heap_bytes, size_t* bytes) {
    AF_REQUIRE(bytes && max_scenarios >= 1 && max_tiles >= 1, AF_E_INVALID,
               "pool_arena_bytes: bad arguments");
    af_domain tmp;
    *bytes = plan_tables(&tmp, max_scenarios, max_tiles) + align_up(heap_bytes, 256);
    return AF_OK;
}

int af_pool_scenario_bytes(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                           size_t* bytes, int32_t* tiles) {
    AF_REQUIRE(desc && bytes, AF_E_INVALID, "scenario_bytes: bad arguments");
    af_domain tmp;
    int rc = configure(&tmp, precision, tuning);
    if (rc) return rc;
    rc = validate(desc);
    if (rc) return rc;
    tmp.min_rows = kMinRowsAlone;  // the most tiles an admission can choose
    HostScn h;
    *bytes = scenario_bytes(&tmp, desc, h);
    if (tiles) *tiles = h.ntiles;
    return AF_OK;
}

int af_pool_create(int precision, const af_tuning* tuning, int max_scenarios, int max_tiles,
                   void* arena, size_t arena_bytes, af_domain** out) {
    AF_REQUIRE(out, AF_E_INVALID, "pool_create: null out");
    af_domain* D = new af_domain();
    int rc = pool_init(D, precision, tuning, max_scenarios, max_tiles, arena, arena_bytes);
    D->pool = true;
    if (rc) {
        delete D;
        return rc;
    }
    *out = D;
    return AF_OK;
}

int af_pool_admit(af_domain* D, const af_scenario_desc* desc, af_stream_t stream, int32_t* slot) {
    AF_REQUIRE(D && desc && slot, AF_E_INVALID, "admit: bad arguments");
    int s = -1;
    int rc = pool_admit(D, desc, joined(D, stream), true, &s);
    if (rc) return rc;
    *slot = s;
    return AF_OK;
}

int af_pool_stage(const af_scenario_desc* desc, int precision, af_staged** out) {
    AF_REQUIRE(desc && out, AF_E_INVALID, "stage: bad arguments");
    *out = nullptr;
    af_staged* p = new (std::nothrow) af_staged();
    AF_REQUIRE(p, AF_E_CAPACITY, "stage: out of host memory");
    int rc = stage_inputs(desc, precision, p->s);
    if (rc) {
        delete p;
        return rc;
    }
    p->desc = *desc;
    *out = p;
    return AF_OK;
}

int af_pool_admit_staged(af_domain* D, const af_staged* staged, af_stream_t stream,
                         int32_t* slot) {
    AF_REQUIRE(D && staged && slot, AF_E_INVALID, "admit_staged: bad arguments");
    int s = -1;
    int rc = pool_admit(D, &staged->desc, joined(D, stream), true, &s, &staged->s);
    if (rc) return rc;
    *slot = s;
    return AF_OK;
}

void af_staged_free(af_staged* staged) { delete staged; }

int af_pool_release(af_domain* D, int32_t slot, af_stream_t stream) {
    AF_REQUIRE(D && slot >= 0 && slot < D->max_scn && D->scn[slot].active, AF_E_INVALID,
               "release: slot %d is not in use", slot);
    HostScn& h = D->scn[slot];
    heap_free(D, h.base, h.bytes);
    h = HostScn{};
    hipStream_t st = joined(D, stream);
    if (D->pool) {
        int rc = retile_on_drain(D, st);
        if (rc) return rc;
    }
    D->dirty = true;
    return AF_OK;
}

int af_pool_poll(af_domain* D, int32_t* finished, int32_t cap, int32_t* n_finished,
                 int32_t* n_active, af_stream_t stream) {
    AF_REQUIRE(D && n_finished, AF_E_INVALID, "poll: bad arguments");
    hipStream_t st = joined(D, stream);
    std::vector<Ctl> c;
    int rc = gather_ctl(D, c, st);
    if (rc) return rc;
    int nf = 0;
    for (size_t a = 0; a < D->active.size(); a++) {
        const int s = D->active[a];
        if (slot_finished(D, s, c[a], (int)(scn_k(D, s) & 1))) {
            if (finished && nf < cap) finished[nf] = s;
            nf++;
        }
    }
    *n_finished = nf;
    if (n_active) *n_active = (int)D->active.size();
    return AF_OK;
}

int af_pool_progress(af_domain* D, int32_t* slots, double* t, int64_t* steps, int32_t* status,
                     int32_t* finished, int32_t cap, int32_t* n, af_stream_t stream) {
    AF_REQUIRE(D && n && cap >= 0, AF_E_INVALID, "progress: bad arguments");
    hipStream_t st = joined(D, stream);
    std::vector<Ctl> c;
    int rc = gather_ctl(D, c, st);
    if (rc) return rc;
    const int na = (int)D->active.size();
    for (int a = 0; a < na && a < cap; a++) {
        const int s = D->active[a];
        const int par = (int)(scn_k(D, s) & 1);
        if (slots) slots[a] = s;
        if (t) t[a] = c[a].t[par];
        if (steps) steps[a] = c[a].n_done;
        if (status) status[a] = c[a].status[par];
        if (finished) finished[a] = slot_finished(D, s, c[a], par) ? 1 : 0;
    }
    *n = na;
    return AF_OK;
}

int af_pool_stats(af_domain* D, int32_t* active, int32_t* tiles, size_t* heap_free,
                  int64_t* step_index) {
    AF_REQUIRE(D, AF_E_INVALID, "stats: null pool");
    int na = 0, nt = 0;
    for (const HostScn& h : D->scn) {
        na += h.active ? 1 : 0;
        nt += h.active ? h.ntiles : 0;
    }
    if (active) *active = na;
    if (tiles) *tiles = nt;
    if (heap_free) *heap_free = heap_free_bytes(D);
    if (step_index) *step_index = D->k;
    return AF_OK;
}

int af_domain_variant(af_domain* D, int32_t* variant) {
    AF_REQUIRE(D && variant, AF_E_INVALID, "variant: bad arguments");
    *variant = D->variant + 1;
    return AF_OK;
}

int af_domain_destroy(af_domain* D) {
    if (!D) return AF_OK;
    if (D->lane_st) {
        (void)hipStreamSynchronize(D->lane_st);
        (void)hipEventDestroy(D->ev_fork);
        (void)hipEventDestroy(D->ev_t0);
        (void)hipEventDestroy(D->ev_join);
        (void)hipStreamDestroy(D->lane_st);
    }
    drop_graph(D->graph, D->gexec);
    drop_graph(D->lgraph, D->lgexec);
    delete D;
    return AF_OK;
}

int af_pool_set_lane(af_domain* D, const int32_t* slots, int32_t n, af_stream_t stream) {
    AF_REQUIRE(D && n >= 0 && (n == 0 || slots), AF_E_INVALID, "set_lane: bad arguments");
    AF_REQUIRE(D->k % 6 == 0, AF_E_STATE,
               "set_lane: the pool's step index %lld is not a multiple of 6", (long long)D->k);
    std::vector<int> v;
    for (int i = 0; i < n; i++) {
        const int s = slots[i];
        AF_REQUIRE(s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
                   "set_lane: slot %d is not in use", s);
        if (std::find(v.begin(), v.end(), s) == v.end()) v.push_back(s);
    }
    std::sort(v.begin(), v.end());
    if (v == D->lane_slots) return AF_OK;
    hipStream_t st = joined(D, stream);
    // a scenario leaving or entering the lane stands at phase 0 on both sides
    for (HostScn& h : D->scn) h.in_lane = false;
    for (int s : v) D->scn[s].in_lane = true;
    D->lane_slots.swap(v);
    return rebuild_tiles(D, st);
}

int af_pool_advance_lane(af_domain* D, int64_t steps, int use_graph, af_stream_t stream) {
    AF_REQUIRE(D && steps >= 0 && steps % 6 == 0, AF_E_INVALID,
               "advance_lane: steps must be a non-negative multiple of 6");
    hipStream_t st = to_stream(stream);
    int src = settle(D, st);
    if (src) return src;
    if (D->nlane == 0 || steps == 0) return AF_OK;
    if (!D->lane_st) {
        AF_HIP(hipStreamCreateWithFlags(&D->lane_st, hipStreamNonBlocking));
        AF_HIP(hipEventCreateWithFlags(&D->ev_fork, hipEventDisableTiming));
        AF_HIP(hipEventCreate(&D->ev_t0));
        AF_HIP(hipEventCreate(&D->ev_join));
    }
    if (use_graph && !D->lgexec) {
        int rc = capture_steps(D, D->d_lane, D->nlane, D->graph_steps, &D->lgraph, &D->lgexec);
        if (rc) return rc;
    }
    // the lane starts behind everything enqueued on `st` so far (table
    // uploads, admissions) and runs beside what is enqueued there next
    AF_HIP(hipEventRecord(D->ev_fork, st));
    AF_HIP(hipStreamWaitEvent(D->lane_st, D->ev_fork, 0));
    AF_HIP(hipEventRecord(D->ev_t0, D->lane_st));
    int64_t left = steps;
    while (use_graph && left >= D->graph_steps) {
        AF_HIP(hipGraphLaunch(D->lgexec, D->lane_st));
        left -= D->graph_steps;
    }
    for (int64_t i = 0; i < left; i++) {
        int rc = launch_tiles(D, D->d_lane, D->nlane, (int)(i % 6), D->lane_st);
        if (rc) return rc;
    }
    AF_HIP(hipEventRecord(D->ev_join, D->lane_st));
    D->lane_pending = true;
    D->lane_timed = true;
    return AF_OK;
}

int af_pool_lane_ms(af_domain* D, float* ms) {
    AF_REQUIRE(D && ms, AF_E_INVALID, "lane_ms: bad arguments");
    AF_REQUIRE(D->lane_timed, AF_E_STATE, "lane_ms: no lane work has been enqueued");
    AF_HIP(hipEventSynchronize(D->ev_join));
    AF_HIP(hipEventElapsedTime(ms, D->ev_t0, D->ev_join));
    return AF_OK;
}

int af_domain_advance(af_domain* D, int64_t max_steps, int use_graph, af_stream_t stream) {
    AF_REQUIRE(D && max_steps >= 0, AF_E_INVALID, "advance: bad arguments");
    hipStream_t st = to_stream(stream);
    int src = settle(D, st);
    if (src) return src;
    if (use_graph) {
        int rc = build_graph(D, st);
        if (rc) return rc;
    }
    int64_t left = max_steps;
    while (left > 0) {
        if (use_graph && (D->k % 6) == 0 && left >= D->graph_steps) {
            AF_HIP(hipGraphLaunch(D->gexec, st));
            D->k += D->graph_steps;
            left -= D->graph_steps;
        } else {
            int rc = launch_step(D, (int)(D->k % 6), st);
            if (rc) return rc;
            D->k++;
            left--;
        }
    }
    return AF_OK;
}

int af_batch_finished(af_domain* D, int* all_done, af_stream_t stream) {
    AF_REQUIRE(D && all_done, AF_E_INVALID, "finished: bad arguments");
    hipStream_t st = joined(D, stream);
    std::vector<Ctl> c;
    int rc = gather_ctl(D, c, st);
    if (rc) return rc;
    int done = 1;
    for (size_t a = 0; a < D->active.size(); a++)
        if (!slot_finished(D, D->active[a], c[a], (int)(scn_k(D, D->active[a]) & 1))) done = 0;
    *all_done = done;
    return AF_OK;
}

int af_domain_run(af_domain* D, int64_t check_every, int use_graph, af_stream_t stream) {
    AF_REQUIRE(D && check_every > 0, AF_E_INVALID, "run: bad arguments");
    hipStream_t st = joined(D, stream);
    std::vector<Ctl> c;
    long long prev_done = -1;
    for (;;) {
        int rc = gather_ctl(D, c, st);
        if (rc) return rc;
        int done = 1;
        long long steps = 0;
        double tsum = 0.0;
        for (size_t a = 0; a < D->active.size(); a++) {
            const int par = (int)(scn_k(D, D->active[a]) & 1);
            if (!slot_finished(D, D->active[a], c[a], par)) done = 0;
            steps += c[a].n_done;
            tsum += c[a].t[par];
        }
        if (done) {
            // every scenario ran to its sim_time or stopped; name the first
            // that stopped
            for (size_t a = 0; a < D->active.size(); a++) {
                const int par = (int)(scn_k(D, D->active[a]) & 1);
                if (c[a].status[par] == 0) continue;
                set_error("scenario %d stopped after step %lld at t = %.17g: %s", D->active[a],
                          (long long)c[a].n_done, c[a].t[par], status_text(c[a].status[par]));
                return AF_E_STATE;
            }
            break;
        }
        // a clock that no longer advances (dt == 0 after an overflow) would never end
        if (prev_done >= 0 && steps > prev_done && !(tsum > D->last_tsum)) {
            set_error("simulation clock stalled at t = %g", tsum);
            return AF_E_STATE;
        }
        prev_done = steps;
        D->last_tsum = tsum;
        rc = af_domain_advance(D, check_every, use_graph, stream);
        if (rc) return rc;
    }
    return AF_OK;
}

int af_batch_set_sim_time(af_domain* D, int s, double sim_time, af_stream_t stream) {
    AF_REQUIRE(D && s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
               "set_sim_time: bad arguments");
    AF_REQUIRE(sim_time == sim_time, AF_E_INVALID, "set_sim_time: NaN");
    hipStream_t st = joined(D, stream);
    D->scn[s].dev.sim_time = sim_time;
    // the kernel reads the table on every launch, graph replays included
    AF_HIP(hipMemcpyAsync(&D->d_scns[s].sim_time, &D->scn[s].dev.sim_time, sizeof(double),
                          hipMemcpyHostToDevice, st));
    AF_HIP(hipStreamSynchronize(st));
    return AF_OK;
}

int af_batch_info(af_domain* D, int s, af_run_info* info, af_stream_t stream) {
    AF_REQUIRE(D && info && s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
               "info: bad arguments");
    hipStream_t st = joined(D, stream);
    int rc = finalize_reports(D, s, st);
    if (rc) return rc;
    Ctl c;
    AF_HIP(hipMemcpyAsync(&c, D->scn[s].dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    const long long k = scn_k(D, s);
    const int tpar = (int)(k & 1);  // every launched step writes t[(k+1) & 1]
    info->t = c.t[tpar];
    info->steps_done = c.n_done;
    info->last_dt = c.last_dt;
    info->n_reports = c.n_reports;
    info->finished = slot_finished(D, s, c, tpar) ? 1 : 0;
    info->hmax = fkey_inv_d(c.hmax[(int)(k % 3)]);
    info->status = c.status[tpar];
    info->reserved = 0;
    return AF_OK;
}

int af_domain_info(af_domain* D, af_run_info* info, af_stream_t stream) {
    return af_batch_info(D, 0, info, stream);
}

int af_batch_read(af_domain* D, int s, int which, double* out, af_stream_t stream) {
    AF_REQUIRE(D && out && s >= 0 && s < D->max_scn && D->scn[s].active && which >= 0 &&
                   which <= 4,
               AF_E_INVALID, "read: bad arguments");
    hipStream_t st = joined(D, stream);
    HostScn& h = D->scn[s];
    long long n_done = 0;
    const int par = current_parity(D, s, &n_done, st);
    if (par < 0) return par;
    const size_t es = elem_size(D->precision);
    const int rows = (which == 2 || which == 4) ? h.ny + 1 : h.ny;
    const void* src = (which == 0) ? h.dev.H[par]
                      : (which == 1 || which == 3) ? h.dev.Qx[par] : h.dev.Qy[par];
    std::vector<char> buf((size_t)h.P * rows * es);
    AF_HIP(hipMemcpyAsync(buf.data(), src, buf.size(), hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    const int nx = h.nx, ny = h.ny;
    auto get = [&](long long k) -> double {
        return es == 8 ? reinterpret_cast<const double*>(buf.data())[k]
                       : (double)reinterpret_cast<const float*>(buf.data())[k];
    };
    auto get_old = [&](long long k) -> double {  // q_old = Q * (1/dx) in the state precision
        if (es == 8) return reinterpret_cast<const double*>(buf.data())[k] * h.dev.dxinv;
        return (double)(reinterpret_cast<const float*>(buf.data())[k] * (float)h.dev.dxinv);
    };
    if (which == 0) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) out[i + (long long)j * nx] = get((long long)j * h.P + i);
        return AF_OK;
    }
    const long long PR = nx + 1;
    for (long long k = 0; k < PR * (ny + 1); k++) out[k] = 0.0;
    const bool old = (which >= 3);
    if (which == 1 || which == 3) {
        for (int j = 0; j < ny; j++)
            for (int f = 0; f <= nx; f++) {
                const long long k = (long long)j * h.P + f;
                out[f + j * PR] = old ? get_old(k) : get(k);
            }
    } else {
        for (int g = 0; g <= ny; g++)
            for (int i = 0; i < nx; i++) {
                const long long k = (long long)g * h.P + i;
                out[i + g * PR] = old ? get_old(k) : get(k);
            }
    }
    return AF_OK;
}

int af_batch_read_cells(af_domain* D, int s, const int64_t* idx, int n, double* out,
                        af_stream_t stream) {
    AF_REQUIRE(D && (n == 0 || (idx && out)) && n >= 0 && s >= 0 && s < D->max_scn &&
                   D->scn[s].active,
               AF_E_INVALID, "read_cells: bad arguments");
    hipStream_t st = joined(D, stream);
    HostScn& h = D->scn[s];
    const int par = current_parity(D, s, nullptr, st);
    if (par < 0) return par;
    const size_t es = elem_size(D->precision);
    const long long ncell = (long long)h.nx * h.ny;
    std::vector<char> v((size_t)n * es);
    for (int k = 0; k < n; k++) {
        if (idx[k] < 0 || idx[k] >= ncell) continue;
        const long long j = idx[k] / h.nx, i = idx[k] % h.nx;
        AF_HIP(hipMemcpyAsync(v.data() + (size_t)k * es,
                              (const char*)h.dev.H[par] + (size_t)(j * h.P + i) * es, es,
                              hipMemcpyDeviceToHost, st));
    }
    AF_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < n; k++) {
        if (idx[k] < 0 || idx[k] >= ncell) {
            out[k] = NAN;
            continue;
        }
        out[k] = es == 8 ? reinterpret_cast<const double*>(v.data())[k]
                         : (double)reinterpret_cast<const float*>(v.data())[k];
    }
    return AF_OK;
}

}  // extern "C"

namespace {

// Output snapshot of one scenario (af_batch_snapshot): H as double and, when
// Vx is set, the velocities of ComputeVelocity (kernels.cpp:293-351), for the
// state entering launch k, which lives in buffer ctl->par[k & 1].  The
// velocity is evaluated in double on the state's values, in the operation
// order of the reference (the compat velocity kernel's order).
template <typename T>
__global__ void __launch_bounds__(256)
k_snapshot(const ScnDev* __restrict__ scn, int kpar, double* __restrict__ Ho,
           double* __restrict__ Vx, double* __restrict__ Vy) {
    const ScnDev& S = *scn;
    const int b = S.ctl->par[kpar];
    const T* __restrict__ H = static_cast<const T*>(S.H[b]);
    const Zt<T>* __restrict__ z = static_cast<const Zt<T>*>(S.z);
    const T* __restrict__ Qx = static_cast<const T*>(S.Qx[b]);
    const T* __restrict__ Qy = static_cast<const T*>(S.Qy[b]);
    Zt<T> nod;
    if constexpr (sizeof(T) == 8) nod = S.nodata_t;
    else nod = INT_MIN;
    const double dx = S.dx;
    const long long n = (long long)S.nx * S.ny;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const long long j = c / S.nx, i = c - j * S.nx;
        const long long o = j * S.P + i;
        const double h = (double)H[o];
        Ho[c] = h;
        if (!Vx) continue;
        if (z[o] == nod || h < 0.001) {
            Vx[c] = 0.0;
            Vy[c] = 0.0;
            continue;
        }
        double hflow = h;
        if (hflow < 0.001) hflow = 0.001;
        const double qa = (double)Qx[o], qb = (double)Qx[o + 1];
        const double va = (fabs(qa) > 1e-10) ? qa / dx / hflow : 0.0;
        const double vb = (fabs(qb) > 1e-10) ? qb / dx / hflow : 0.0;
        Vx[c] = (fabs(va) > fabs(vb)) ? va : vb;
        const double qc = (double)Qy[o], qd = (double)Qy[o + S.P];
        const double vc = (fabs(qc) > 1e-10) ? qc / dx / hflow : 0.0;
        const double vd = (fabs(qd) > 1e-10) ? qd / dx / hflow : 0.0;
        Vy[c] = (fabs(vc) > fabs(vd)) ? vc : vd;
    }
}

}  // namespace

extern "C" {

int af_batch_snapshot(af_domain* D, int s, double* H_out, double* Vx_out, double* Vy_out,
                      af_stream_t stream) {
    AF_REQUIRE(D && H_out && s >= 0 && s < D->max_scn && D->scn[s].active &&
                   ((Vx_out == nullptr) == (Vy_out == nullptr)),
               AF_E_INVALID, "snapshot: bad arguments");
    hipStream_t st = joined(D, stream);
    const HostScn& h = D->scn[s];
    const long long n = (long long)h.nx * h.ny;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 16384);
    const int kpar = (int)(scn_k(D, s) & 1);
    if (D->precision == AF_FP64)
        hipLaunchKernelGGL(k_snapshot<double>, dim3(blocks), dim3(256), 0, st, D->d_scns + s,
                           kpar, H_out, Vx_out, Vy_out);
    else
        hipLaunchKernelGGL(k_snapshot<float>, dim3(blocks), dim3(256), 0, st, D->d_scns + s, kpar,
                           H_out, Vx_out, Vy_out);
    AF_HIP(hipGetLastError());
    return AF_OK;
}

int af_domain_read(af_domain* D, int which, double* out, af_stream_t stream) {
    return af_batch_read(D, 0, which, out, stream);
}

int af_batch_read_dt_log(af_domain* D, int s, double* out, int64_t cap, int64_t* n,
                         af_stream_t stream) {
    AF_REQUIRE(D && n && s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
               "dt_log: bad arguments");
    hipStream_t st = joined(D, stream);
    long long done = 0;
    if (current_parity(D, s, &done, st) < 0) return AF_E_HIP;
    long long avail = std::min<long long>(done, D->dt_log_cap);
    long long m = std::min<long long>(avail, cap);
    if (m > 0 && out)
        AF_HIP(hipMemcpy(out, D->scn[s].dev.dt_log, (size_t)m * 8, hipMemcpyDeviceToHost));
    *n = avail;
    return AF_OK;
}

int af_domain_read_dt_log(af_domain* D, double* out, int64_t cap, int64_t* n,
                          af_stream_t stream) {
    return af_batch_read_dt_log(D, 0, out, cap, n, stream);
}

int af_batch_read_reports(af_domain* D, int s, double* out, int32_t cap, int32_t* n,
                          af_stream_t stream) {
    AF_REQUIRE(D && n && s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
               "reports: bad arguments");
    hipStream_t st = joined(D, stream);
    int rc = finalize_reports(D, s, st);
    if (rc) return rc;
    Ctl c;
    AF_HIP(hipMemcpyAsync(&c, D->scn[s].dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    int avail = std::min(c.n_reports, D->report_cap);
    int m = std::min(avail, cap);
    if (m > 0 && out)
        AF_HIP(hipMemcpy(out, D->scn[s].dev.reports, (size_t)m * 32, hipMemcpyDeviceToHost));
    *n = c.n_reports;
    return AF_OK;
}

int af_domain_read_reports(af_domain* D, double* out, int32_t cap, int32_t* n,
                           af_stream_t stream) {
    return af_batch_read_reports(D, 0, out, cap, n, stream);
}

int af_batch_volume(af_domain* D, int s, double* vol, double* area, af_stream_t stream) {
    AF_REQUIRE(D && s >= 0 && s < D->max_scn && D->scn[s].active, AF_E_INVALID,
               "volume: bad arguments");
    hipStream_t st = joined(D, stream);
    HostScn& h = D->scn[s];
    const int par = current_parity(D, s, nullptr, st);
    if (par < 0) return par;
    const long long n = (long long)h.nx * h.ny;
    unsigned blocks = blocks_for(n, 256);
    if (blocks > (unsigned)kVolBlocks) blocks = kVolBlocks;
    if (D->precision == AF_FP64)
        hipLaunchKernelGGL(k_volume<double>, dim3(blocks), dim3(256), 0, st,
                           (const double*)h.dev.H[par], h.P, h.nx, h.ny, h.dev.dA, D->d_volbuf);
    else
        hipLaunchKernelGGL(k_volume<float>, dim3(blocks), dim3(256), 0, st,
                           (const float*)h.dev.H[par], h.P, h.nx, h.ny, h.dev.dA, D->d_volbuf);
    AF_HIP(hipGetLastError());
    std::vector<double> part(3 * blocks);
    AF_HIP(hipMemcpyAsync(part.data(), D->d_volbuf, part.size() * 8, hipMemcpyDeviceToHost, st));
    AF_HIP(hipStreamSynchronize(st));
    long double s_ = 0.0L, w = 0.0L;
    for (unsigned b = 0; b < blocks; b++) {
        s_ += (long double)part[3 * b] + (long double)part[3 * b + 1];
        w += part[3 * b + 2];
    }
    if (vol) *vol = (double)s_;
    if (area) *area = (double)(w * h.dev.dA);
    return AF_OK;
}

int af_domain_volume(af_domain* D, double* vol, double* area, af_stream_t stream) {
    return af_batch_volume(D, 0, vol, area, stream);
}

int af_domain_state_ptr(af_domain* D, int which, void** ptr, int64_t* pitch_elems) {
    AF_REQUIRE(D && ptr && which >= 0 && which <= 2, AF_E_INVALID, "state_ptr: bad arguments");
    long long done = 0;
    int par = current_parity(D, 0, &done, nullptr);
    if (par < 0) return par;
    HostScn& h = D->scn[0];
    *ptr = which == 0 ? h.dev.H[par] : which == 1 ? h.dev.Qx[par] : h.dev.Qy[par];
    if (pitch_elems) *pitch_elems = h.P;
    return AF_OK;
}

int af_batch_size(af_domain* D, int* n_scenarios, int* n_tiles) {
    AF_REQUIRE(D, AF_E_INVALID, "batch_size: null");
    int na = 0;
    for (const HostScn& h : D->scn) na += h.active ? 1 : 0;
    if (n_scenarios) *n_scenarios = na;
    if (n_tiles) *n_tiles = D->ntiles;
    return AF_OK;
}

int af_batch_launch_step(af_domain* D, af_stream_t stream) {
    AF_REQUIRE(D, AF_E_INVALID, "launch_step: null");
    int rc = settle(D, to_stream(stream));
    if (rc) return rc;
    rc = launch_step(D, (int)(D->k % 6), to_stream(stream));
    if (rc) return rc;
    D->k++;
    return AF_OK;
}

}  // extern "C"
