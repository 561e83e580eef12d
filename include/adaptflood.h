/*
 * adaptflood.h -- C ABI of the MI355X-native local-inertial flood step.
 *
 * Two layers share one set of device arithmetic:
 *
 *  1. The reference call surface (FP64, reference memory layout).  Each
 *     af_<Name>_GPU replaces the reference's <Name>_GPU C++ function of the
 *     same name (src/kernels.hpp:8-37; the point-source injection has no name
 *     in the reference and lives inline at src/main.cpp:244-262).  A Kokkos
 *     `RealView&` becomes a `double*` device pointer, an `IntView&` becomes an
 *     `int32_t*` device pointer, and every call takes the HIP stream it is
 *     enqueued on (NULL = the default stream).  Layout as in the reference:
 *     H and DEM hold xsz*ysz cells, idx = i + j*xsz, row 0 is the north row;
 *     Qx, Qy, Qxold, Qyold hold (xsz+1)*(ysz+1) faces with pitch xsz+1.
 *     Like the reference, each call returns only after its work is done
 *     (it synchronises the stream).
 *
 *  2. The fused production path (FP32 or FP64 state): a domain owns its state
 *     in a caller-provided device arena and advances it by one fused HIP
 *     kernel per step (fluxes, line and point boundaries, rain, depth update,
 *     overrides and the next step's CFL maximum in one pass), optionally
 *     replayed from a captured HIP graph.  Batched ensembles of ragged
 *     domains advance together in one launch per step.
 *
 * Every entry point returns an int status: AF_OK (0) or a negative AF_E_*
 * code; af_last_error() gives a message for the calling thread.  The library
 * never allocates device memory: the caller owns every device buffer.  The
 * one exception is a 256-byte static device array per GPU, the scratch of
 * af_FloodplainQ_GPU_noalloc, which the loader places with the code.
 *
 * Threads: the per-function entry points keep no state between calls and
 * may be called from several threads at once (each with its own buffers and
 * FloodplainQ workspace), except af_FloodplainQ_GPU_noalloc, whose calls on
 * one GPU share its scratch and must not overlap.  A domain, batch or pool
 * (af_domain) is not thread-safe: calls on one af_domain must not overlap.
 */
#ifndef ADAPTFLOOD_AMD_H
#define ADAPTFLOOD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* af_stream_t; /* a hipStream_t */

enum {
    AF_OK = 0,
    AF_E_INVALID = -1,  /* bad argument (size, pointer, precision, index) */
    AF_E_HIP = -2,      /* a HIP runtime call failed */
    AF_E_IO = -3,       /* a file could not be opened or parsed */
    AF_E_CAPACITY = -4, /* arena or table too small */
    AF_E_STATE = -5     /* call out of order (e.g. run before load) */
};

const char* af_last_error(void);
const char* af_version(void);

/* ---------------------------------------------------------------------- */
/* 1. Reference call surface (FP64, reference layout)                      */
/* ---------------------------------------------------------------------- */

/* Bytes of device workspace af_FloodplainQ_GPU needs for its H-max
 * reduction (the reference lets Kokkos allocate this implicitly). */
#define AF_FLOODPLAINQ_WORKSPACE_BYTES 256

/* Replaces UpdateQs_GPU (src/kernels.hpp:8-9, src/kernels.cpp:6-19):
 * Qxold = Qx * (1/dx), Qyold = Qy * (1/dx) over all (xsz+1)*(ysz+1) faces. */
int af_UpdateQs_GPU(const double* Qx, const double* Qy, double* Qxold, double* Qyold, int xsz,
                    int ysz, double dx, af_stream_t stream);

/* Replaces FloodplainQ_GPU (src/kernels.hpp:12-16, src/kernels.cpp:22-146):
 * H-max over cells whose DEM != nodata, dt = Hmax > DepthThresh ?
 * cfl*dx/sqrt(g*Hmax) : InitTstep written to *Tstep (host), then the interior
 * x- and y-face inertial fluxes.  dy, dhlin, Qlimfact and dA are accepted
 * and unused, as in the reference.  `workspace` is device memory of at least
 * AF_FLOODPLAINQ_WORKSPACE_BYTES. */
int af_FloodplainQ_GPU(const double* H, const double* DEM, double* Qx, double* Qy,
                       const double* Qxold, const double* Qyold, int xsz, int ysz, double dx,
                       double dy, double FPn, double DepthThresh, double MaxHflow, double dhlin,
                       double* Tstep, double Qlimfact, double dA, double cfl, double g,
                       double InitTstep, double theta, double nodata, void* workspace,
                       af_stream_t stream);

/* af_FloodplainQ_GPU without the workspace argument, and allocating nothing:
 * no hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture elsewhere in their loop, serving loops, harnesses
 * that forbid allocation).  The H-max key lives in a static device array of
 * AF_FLOODPLAINQ_WORKSPACE_BYTES (256 bytes) that each GPU holds until the
 * process ends.  The scratch need is one 64-bit key whatever the grid, so
 * every grid the other arguments can describe fits it.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two threads, or two
 * streams) race and may return each other's Tstep.  Each call clears the key
 * before it reduces into it.  Results equal af_FloodplainQ_GPU's bitwise. */
int af_FloodplainQ_GPU_noalloc(const double* H, const double* DEM, double* Qx, double* Qy,
                               const double* Qxold, const double* Qyold, int xsz, int ysz,
                               double dx, double dy, double FPn, double DepthThresh,
                               double MaxHflow, double dhlin, double* Tstep, double Qlimfact,
                               double dA, double cfl, double g, double InitTstep, double theta,
                               double nodata, af_stream_t stream);

/* Replaces BCs_GPU (src/kernels.hpp:19-21, src/kernels.cpp:149-256): FREE or
 * CLOSED edge faces from the pre-update H.  dt is unused, as in the reference. */
int af_BCs_GPU(double* Qx, double* Qy, const double* H, const double* DEM, int xsz, int ysz,
               double dx, double g, double dt, double nodata, int west_free, int east_free,
               int north_free, int south_free, af_stream_t stream);

/* Replaces ApplyPointBCs_GPU (src/kernels.hpp:24-25, src/kernels.cpp:259-277):
 * bc_type 1 (FREE) sets H = 0, 2/3 (HFIX/HVAR) set H = bc_value.  Points that
 * share a cell resolve in array order (the last one wins).  ncols, dx, dA and
 * Tstep are unused, as in the reference.  Every bc_idx must lie in
 * [0, xsz*ysz): like the reference, the call does not know the grid size.
 * Up to 8192 points are grouped by cell with one sort in LDS (O(n log^2 n)
 * work in one workgroup); longer lists fall back to a scan of the list for
 * each point (O(num_bcs^2) work). */
int af_ApplyPointBCs_GPU(double* H, const int32_t* bc_idx, const int32_t* bc_type,
                         const double* bc_value, int num_bcs, int ncols, double dx, double dA,
                         double Tstep, af_stream_t stream);

/* Replaces UpdateH_GPU (src/kernels.hpp:28, src/kernels.cpp:280-290):
 * H += Tstep*(Qx_in - Qx_out + Qy_in - Qy_out)/dA, clamped at 0, all cells. */
int af_UpdateH_GPU(double* H, const double* Qx, const double* Qy, int xsz, int ysz, double dA,
                   double Tstep, af_stream_t stream);

/* Replaces Rainfall_GPU (src/kernels.hpp:36-37, src/kernels.cpp:354-367):
 * if rain_rate > 0, H += rain_rate*Tstep on cells whose DEM != nodata. */
int af_Rainfall_GPU(double* H, const double* DEM, int xsz, int ysz, double rain_rate,
                    double Tstep, double nodata, af_stream_t stream);

/* Replaces the inline AddFlowBCs kernel (src/main.cpp:256-260):
 * H[ps_idx[s]] += ps_flow[s], clamped at 0.  Sources that share a cell are
 * applied in array order.  Every ps_idx must lie in [0, xsz*ysz).  Up to
 * 8192 sources are grouped by cell with one sort in LDS; longer lists fall
 * back to a scan of the list for each source (O(num_flow_bcs^2) work). */
int af_AddFlowBCs_GPU(double* H, const int32_t* ps_idx, const double* ps_flow, int num_flow_bcs,
                      af_stream_t stream);

/* Replaces ComputeVelocity_GPU (src/kernels.hpp:31-33, src/kernels.cpp:293-351). */
int af_ComputeVelocity_GPU(double* Vx_out, double* Vy_out, const double* H, const double* DEM,
                           const double* Qx, const double* Qy, int xsz, int ysz, double dx,
                           double nodata, af_stream_t stream);

/* ---------------------------------------------------------------------- */
/* 2. Fused path                                                           */
/* ---------------------------------------------------------------------- */

enum { AF_FP32 = 0, AF_FP64 = 1 };

/* Point-boundary type codes of src/types.hpp:41-48 */
enum { AF_BC_CLOSED = 0, AF_BC_FREE = 1, AF_BC_HFIX = 2, AF_BC_HVAR = 3, AF_BC_QFIX = 4,
       AF_BC_QVAR = 5 };

/* Everything one scenario needs, as host arrays.  Mirrors what main.cpp
 * builds before its time loop (src/main.cpp:50-127, 166-194). */
typedef struct af_scenario_desc {
    int32_t nx, ny;          /* ncols, nrows */
    double dx;               /* cellsize (dx = dy, dA = dx*dx; main.cpp:115) */
    double nodata;           /* NODATA_value of the DEM */
    const double* dem;       /* nx*ny, idx = i + j*nx, row 0 north */
    double fpn;              /* Manning n (par.fpn) */
    double depth_thresh;     /* par.depth_thresh */
    double max_hflow;        /* 10.0 (main.cpp:118) */
    double init_tstep;       /* par.init_tstep */
    double cfl;              /* par.cfl */
    double g;                /* 9.81 (main.cpp:123) */
    double theta;            /* 1.0 (main.cpp:124); the fused path requires 1.0 */
    double sim_time;         /* par.sim_time */
    double mass_int;         /* par.mass_int (device-side mass reports) */
    int32_t west_free, east_free, north_free, south_free; /* main.cpp:62-70 */
    /* point boundaries in file order, any type (AF_BC_*); x/y are grid
     * coordinates as read_bci computes them (src/io.cpp:185-187). */
    int32_t n_points;
    const int32_t* pt_x;
    const int32_t* pt_y;
    const int32_t* pt_type;
    const double* pt_value;  /* QFIX/HFIX value; value used when a series is empty */
    const int32_t* pt_series_off; /* offset into series_t/series_v (QVAR/HVAR) */
    const int32_t* pt_series_len; /* 0 = empty series */
    int32_t n_series;        /* total series entries */
    const double* series_t;
    const double* series_v;
    /* uniform rain (m/s); rain_enabled follows read_rain (src/io.cpp:123-124) */
    int32_t rain_enabled;
    int32_t n_rain;
    const double* rain_t;
    const double* rain_r;
    /* initial state (NULL = zeros, as main.cpp:156-160) in the reference layout */
    const double* h0;        /* nx*ny */
} af_scenario_desc;

/* Scenario status bits (af_run_info.status).  A scenario with a non-zero
 * status has stopped for good: its tiles skip every later step, and it
 * counts as finished.  The reference has no such stop: its H max skips NaN
 * (src/kernels.cpp:30-33), so a NaN depth runs to the end unnoticed, and an
 * infinite depth gives dt = 0 and a loop that never ends (main.cpp:227). */
enum {
    AF_SCN_OK = 0,
    AF_SCN_NONFINITE = 1,    /* a step produced a NaN or infinite depth */
    AF_SCN_STALLED = 2       /* t + dt == t: the clock could not advance */
};

typedef struct af_run_info {
    double t;                /* simulated time after the last executed step */
    int64_t steps_done;      /* steps executed since af_domain_create */
    double last_dt;          /* Tstep of the last executed step */
    int32_t n_reports;       /* mass reports recorded so far */
    int32_t finished;        /* t >= sim_time, or stopped (status != 0) */
    double hmax;             /* max H over valid cells of the current state */
    int32_t status;          /* AF_SCN_* bits */
    int32_t reserved;
} af_run_info;

typedef struct af_domain af_domain; /* opaque, host side */

/* Tile shape: each wavefront owns a strip of 64*V columns (V = 4 at FP32,
 * 2 at FP64) and marches down `tile_rows` rows.  0 = library default: 64 rows
 * at FP32 and 32 at FP64, halved per scenario while the scenario would offer
 * fewer than 2048 wavefronts, down to 2 rows for a scenario that runs alone
 * (a domain, a batch of one, or an admission to a pool holding fewer than
 * 2048 tiles) and to 8 rows otherwise; tiles under 8 rows only while the
 * scenario stays under 2048 tiles, under 4 rows only while at most 512.  A
 * pool that drains below 2048 tiles moves its scenarios to the height they
 * would take alone.  No field changes any result: the
 * tile shape, variant and order only change how the work is split. */
typedef struct af_tuning {
    int32_t tile_rows;
    int32_t waves_per_block; /* 1 or 4 */
    int32_t graph_steps;     /* steps per captured graph (multiple of 6); 0 = default */
    int32_t dt_log_cap;      /* device dt log entries kept (0 = none) */
    int32_t report_cap;      /* device mass-report records kept */
    int32_t variant;         /* 0 = library default; 1..3 = step-kernel variant: 1 one-row
                              * loop, 2 and 3 two-row bursts at lower target occupancy */
    int32_t tile_order;      /* order of the launch's tile table: 0 = by scenario, segment,
                              * strip; 1 = largest tiles (rows x columns) first, so that the
                              * last workgroups dispatched are the short ones */
} af_tuning;

/* Device bytes the domain needs for `desc` at `precision`. */
int af_domain_arena_bytes(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                          size_t* bytes);

/* Builds a domain in `arena` (device memory of at least the queried size,
 * 256-byte aligned) and uploads the scenario.  The arena stays owned by the
 * caller and must outlive the domain. */
int af_domain_create(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                     void* arena, size_t arena_bytes, af_stream_t stream, af_domain** out);
int af_domain_destroy(af_domain* d);

/* Enqueues up to `max_steps` fused steps (each one a no-op once t >= sim_time).
 * use_graph != 0 replays a captured graph of graph_steps steps. */
int af_domain_advance(af_domain* d, int64_t max_steps, int use_graph, af_stream_t stream);
/* Runs until t >= sim_time; checks completion every `check_every` steps. */
int af_domain_run(af_domain* d, int64_t check_every, int use_graph, af_stream_t stream);
/* Synchronises `stream` and reads the run counters. */
int af_domain_info(af_domain* d, af_run_info* info, af_stream_t stream);

/* Copies the current state out in the reference layout, converted to double:
 * which = 0: H (nx*ny); 1: Qx; 2: Qy ((nx+1)*(ny+1), pitch nx+1, the
 * padding entries zero); 3: Qxold; 4: Qyold.  `out` is host memory. */
int af_domain_read(af_domain* d, int which, double* out, af_stream_t stream);
/* Copies up to `cap` dt values of executed steps (from the device log). */
int af_domain_read_dt_log(af_domain* d, double* out, int64_t cap, int64_t* n, af_stream_t stream);
/* Mass reports as (t, Tstep, Vol, Area) quadruples, main.cpp:290-304. */
int af_domain_read_reports(af_domain* d, double* out, int32_t cap, int32_t* n,
                           af_stream_t stream);
/* Device volume sum, FP64 with compensation: sum over all cells of H*dA. */
int af_domain_volume(af_domain* d, double* vol, double* area, af_stream_t stream);
/* Device pointer and pitch of the current H (for zero-copy consumers). */
int af_domain_state_ptr(af_domain* d, int which, void** ptr, int64_t* pitch_elems);

/* Batches: n ragged scenarios (each with its own grid, physics, points, rain,
 * dt and clock) in one arena, advanced together by one launch per step over a
 * table of tiles.  The af_domain_* calls above act on scenario 0 of a batch;
 * the af_batch_* calls take the scenario index. */
int af_batch_arena_bytes(const af_scenario_desc* descs, int n, int precision,
                         const af_tuning* tuning, size_t* bytes);
int af_batch_create(const af_scenario_desc* descs, int n, int precision, const af_tuning* tuning,
                    void* arena, size_t arena_bytes, af_stream_t stream, af_domain** out);
int af_batch_size(af_domain* d, int* n_scenarios, int* n_tiles);
/* *all_done = 1 when every scenario has reached its sim_time. */
int af_batch_finished(af_domain* d, int* all_done, af_stream_t stream);
int af_batch_info(af_domain* d, int scenario, af_run_info* info, af_stream_t stream);
/* Moves the time at which `scenario` stops (the loop condition t < sim_time
 * of src/main.cpp:227).  A caller that stops at each save time, writes its
 * output and moves the stop time on reproduces the reference's save cadence
 * (main.cpp:327): the run stops at the first step with t >= the stop time. */
int af_batch_set_sim_time(af_domain* d, int scenario, double sim_time, af_stream_t stream);
/* H at n cells given as reference indices i + j*nx (the stage gauges of
 * main.cpp:306-322); indices outside the grid give NaN.  `out` is host memory. */
int af_batch_read_cells(af_domain* d, int scenario, const int64_t* idx, int n, double* out,
                        af_stream_t stream);
int af_batch_read(af_domain* d, int scenario, int which, double* out, af_stream_t stream);
/* Output snapshot of `scenario` on the device, replacing the host copy of H
 * and the ComputeVelocity_GPU call at a save (src/main.cpp:326-350; velocity
 * of src/kernels.cpp:293-351): H as double in the reference layout (nx*ny,
 * i + j*nx, north row first) into the device buffer H_out, and, when Vx_out
 * and Vy_out are both non-NULL, the velocities likewise.  It is enqueued on
 * `stream` behind the steps already enqueued there and reads the state those
 * steps leave; it does not wait for the host, so further steps may be
 * enqueued at once and the buffers copied out after an event on `stream`. */
int af_batch_snapshot(af_domain* d, int scenario, double* H_out, double* Vx_out, double* Vy_out,
                      af_stream_t stream);
int af_batch_read_dt_log(af_domain* d, int scenario, double* out, int64_t cap, int64_t* n,
                         af_stream_t stream);
int af_batch_read_reports(af_domain* d, int scenario, double* out, int32_t cap, int32_t* n,
                          af_stream_t stream);
int af_batch_volume(af_domain* d, int scenario, double* vol, double* area, af_stream_t stream);
/* Enqueues exactly one fused step without a graph (for timing one launch). */
int af_batch_launch_step(af_domain* d, af_stream_t stream);

/* Pools: one fixed arena (a table region, then a heap of scenario extents)
 * into which scenarios are admitted and from which they are released while
 * the others keep stepping: continuous batching.  A pool is an af_domain;
 * every af_domain_* / af_batch_* call applies to it, with slot indices as
 * scenario indices.  Admission and release take effect between launches. */
int af_pool_arena_bytes(int max_scenarios, int max_tiles, size_t heap_bytes, size_t* bytes);
/* Heap bytes and tile-table entries one scenario needs (its HBM footprint):
 * with the default tiling, the most an admission can take (its shortest
 * tiles). */
int af_pool_scenario_bytes(const af_scenario_desc* desc, int precision, const af_tuning* tuning,
                           size_t* bytes, int32_t* tiles);
int af_pool_create(int precision, const af_tuning* tuning, int max_scenarios, int max_tiles,
                   void* arena, size_t arena_bytes, af_domain** out);
/* AF_E_CAPACITY when no slot, tile-table room or heap extent is free. */
int af_pool_admit(af_domain* pool, const af_scenario_desc* desc, af_stream_t stream,
                  int32_t* slot);
/* Admission in two parts, for a server that admits while the pool steps:
 * af_pool_stage does the host-only part (validation, and the terrain and
 * initial depths converted to the device format of `precision`, in pinned
 * host memory that is cached for reuse, never returned to the system while
 * the process runs); it touches no pool and launches nothing, so it may run
 * on any thread, while launches are in flight.  af_pool_admit_staged then does the rest, as af_pool_admit does
 * (AF_E_CAPACITY when there is no room; the staged scenario may be admitted
 * later, or into another pool of the same precision).  The arrays of `desc`
 * must stay valid until the staged scenario is freed.  No counterpart in the
 * reference, whose main.cpp:166-226 reads and uploads one scenario per run. */
typedef struct af_staged af_staged; /* opaque, host side */
int af_pool_stage(const af_scenario_desc* desc, int precision, af_staged** out);
int af_pool_admit_staged(af_domain* pool, const af_staged* staged, af_stream_t stream,
                         int32_t* slot);
void af_staged_free(af_staged* staged);
int af_pool_release(af_domain* pool, int32_t slot, af_stream_t stream);
/* Slots whose clock reached their sim_time or that stopped with a non-zero
 * status (up to cap), and the active count. */
int af_pool_poll(af_domain* pool, int32_t* finished, int32_t cap, int32_t* n_finished,
                 int32_t* n_active, af_stream_t stream);
/* Every active slot's progress, read with one gather launch, one copy and one
 * synchronisation: slot index, t, executed steps, status (AF_SCN_*) and
 * whether it is finished (t >= sim_time or status != 0).  Entry a of each
 * non-null array describes the a-th active slot, for a < min(cap, *n);
 * *n is the number of active slots. */
int af_pool_progress(af_domain* pool, int32_t* slots, double* t, int64_t* steps, int32_t* status,
                     int32_t* finished, int32_t cap, int32_t* n, af_stream_t stream);
/* The lane: the listed slots' tiles move from the pool's tile table to a
 * second one (an empty list empties it; released slots leave it), at a pool
 * step index that is a multiple of 6.  af_pool_advance_lane enqueues `steps`
 * steps (a multiple of 6) of the lane's scenarios on a stream of the pool's
 * own, from a captured graph of graph_steps launches when use_graph is set:
 * they start behind the work enqueued on `stream` so far and run beside the
 * af_domain_advance launches enqueued there next, which no longer step them.
 * Every other call on the pool first makes `stream` wait for that lane work.
 * A scenario that needs many more steps than the others (a small time step)
 * so takes extra steps at the cost of its own small launches, overlapped with
 * the pool's.  Each scenario keeps its own dt and clock, and its results are
 * those it gives alone (the phase of its rotating slots is kept: it enters
 * and leaves the lane at phase 0 and steps there in multiples of 6).
 * af_pool_lane_ms: duration of the last lane work (waits for it to end).
 * No counterpart in the reference, which runs one scenario per process. */
int af_pool_set_lane(af_domain* pool, const int32_t* slots, int32_t n, af_stream_t stream);
int af_pool_advance_lane(af_domain* pool, int64_t steps, int use_graph, af_stream_t stream);
int af_pool_lane_ms(af_domain* pool, float* ms);
int af_pool_stats(af_domain* pool, int32_t* active, int32_t* tiles, size_t* heap_free,
                  int64_t* step_index);

/* The kernel variant the next launches use (1..3, af_tuning.variant's
 * numbering).  With af_tuning.variant 0 at FP64 it follows the flow: the
 * two-row variant (2) while the largest scenario has 8192 tiles or more and
 * under 3% of the cells are wet, the one-row variant (1) otherwise; the wet
 * count (the last mass report's) is re-read whenever every control block is
 * read back (af_domain_run, af_batch_finished, af_pool_poll).  No counterpart
 * in the reference (it has one kernel sequence). */
int af_domain_variant(af_domain* d, int32_t* variant);

/* ---------------------------------------------------------------------- */
/* Input formats (host): .par, .dem.ascii, .bci, .bdy, rain                */
/* ---------------------------------------------------------------------- */

typedef struct af_par {
    char dem_file[512], bci_file[512], bdy_file[512], rain_file[512], stage_file[512],
        output_dir[512];
    double sim_time, save_int, mass_int, stage_int, init_tstep, cfl, fpn, depth_thresh, nodata;
    int32_t output_format, voutput, async_output;
} af_par;

typedef struct af_case af_case; /* opaque parsed case */

/* Parameters with the defaults of src/types.hpp:18-38. */
void af_par_defaults(af_par* p);
/* Replaces read_par (src/io.cpp:9-47); keys as the code reads them.  `p` is
 * first set to the defaults (af_par_defaults), then overwritten by the keys the
 * file holds; on an error its contents are unspecified. */
int af_read_par(const char* path, af_par* p);

af_case* af_case_new(void);
void af_case_free(af_case* c);
/* Replaces read_dem (src/io.cpp:132-161). */
int af_case_read_dem(af_case* c, const char* path);
/* Replaces read_bci (src/io.cpp:164-254); needs the DEM header first. */
int af_case_read_bci(af_case* c, const char* path);
/* Replaces read_bdy (src/io.cpp:257-296). */
int af_case_read_bdy(af_case* c, const char* path);
/* Replaces read_rain (src/io.cpp:103-129). */
int af_case_read_rain(af_case* c, const char* path);
/* Replaces read_stage_file (src/io.cpp:50-85); needs the DEM header first. */
int af_case_read_stage(af_case* c, const char* path);
/* Stage gauge i: grid x, y and name (cap bytes); *n = number of gauges
 * (i < 0 only counts). */
int af_case_stage(af_case* c, int32_t i, int32_t* grid_x, int32_t* grid_y, char* name,
                  int32_t cap, int32_t* n);
/* Runs the input sequence of main.cpp:24-113 for a .par file: par, DEM, BCI,
 * BDY (all required), the optional rain file and the optional stage file. */
int af_case_load(af_case* c, const char* par_path);
/* Loads the inputs a filled parameter set names (DEM, BCI, BDY, and the rain
 * and stage files when set), as af_case_load does after reading its .par
 * file.  main.cpp:26-31 runs this way on the defaults (af_par_defaults) when
 * no .par file is given. */
int af_case_load_par(af_case* c, const af_par* par);
/* Fills `desc` with pointers into the case (valid while the case lives),
 * with the constants of main.cpp:115-127. */
int af_case_scenario(af_case* c, af_scenario_desc* desc);
int af_case_par(af_case* c, af_par* p);
/* DEM header: ncols, nrows, xll, yll, cellsize, nodata. */
int af_case_dem_header(af_case* c, int32_t* ncols, int32_t* nrows, double* xll, double* yll,
                       double* cellsize, double* nodata);
/* Line boundaries found (count of FREE sides and their letters, in file order). */
int af_case_line_bcs(af_case* c, char* sides, int32_t cap, int32_t* n);
/* Point i: grid x, y, type, value, series length, name (cap bytes). */
int af_case_point(af_case* c, int32_t i, int32_t* x, int32_t* y, int32_t* type, double* value,
                  int32_t* series_len, char* name, int32_t cap);
int af_case_point_series(af_case* c, int32_t i, double* times, double* values, int32_t cap);
/* Rain series (rate m/s, time s). */
int af_case_rain(af_case* c, int32_t* enabled, double* times, double* rates, int32_t cap,
                 int32_t* n);
/* Linear interpolation of src/io.cpp:299-311 (host). */
double af_interp_series(const double* times, const double* values, int32_t n, double empty_value,
                        double t);
/* ESRI-ASCII writer in the format of write_asc (src/io.cpp:314-334). */
int af_write_asc(const char* path, const double* data, int32_t ncols, int32_t nrows, double xll,
                 double yll, double cellsize);
/* NetCDF output of the asynchronous writer (src/async_output.cpp:18-54
 * creates the file, :108-141 appends one save): dimensions time (unlimited),
 * x, y; variables time (double, seconds), H(time, y, x) (float, units "m",
 * long_name "water_depth", _FillValue = (float)nodata) and, with velocity,
 * Vx and Vy (float, "m/s"); global xllcorner, yllcorner, cellsize.  The file
 * is in the classic 64-bit-offset NetCDF format (the reference writes
 * NetCDF-4 through libnetcdf); append converts the double grids to float as
 * the reference does and updates the record count, so the file is whole
 * after every call. */
typedef struct af_nc af_nc;
int af_nc_create(const char* path, int32_t ncols, int32_t nrows, double xll, double yll,
                 double cellsize, double nodata, int32_t velocity, af_nc** out);
int af_nc_append(af_nc* nc, double t, const double* H, const double* Vx, const double* Vy);
int af_nc_close(af_nc* nc);

/* ESRI-ASCII writer in the format of the asynchronous writer
 * (AsyncOutputManager::write_asc_file, src/async_output.cpp:158-176):
 * corners "%.6f", NODATA_value the grid's own value as "%.0f". */
int af_write_asc_async(const char* path, const double* data, int32_t ncols, int32_t nrows,
                       double xll, double yll, double cellsize, double nodata);

#ifdef __cplusplus
}
#endif

#endif /* ADAPTFLOOD_AMD_H */
