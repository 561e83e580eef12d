// driver.cpp -- lisflood_amd: the reference program (src/main.cpp) over the
// MI355X library.
//
//   lisflood_amd [--path compat|fp64|fp32] [--quiet] [parfile]
//
// Without a parameter file it runs on the defaults, reading dem.ascii,
// bci.bci and bdy.bdy from the working directory (main.cpp:28-31).
// --path compat (default): the reference's own call sequence
//   (main.cpp:227-288) through the drop-in wrappers of adaptflood_compat.hpp,
//   FP64 on the reference memory layout, one synchronous call per kernel.
// --path fp64 | fp32: the fused one-launch-per-step domain, replayed from
//   HIP graphs; it stops at each save time (the first step with
//   t >= next_save, as main.cpp:327 saves) to write the output.
//
// Console lines and output files follow main.cpp: the mass report line
// "t=%.0fs: Tstep=%.2f, Vol=%.0fm3, Area=%.0fm2" (main.cpp:300), H_%04d.asc
// at t = 0 and at every save, H_final.asc, Vx_final.asc / Vy_final.asc when
// the .par file asks for velocity output (main.cpp:380-392), and stage.csv
// when it names a stage file (main.cpp:95-113, 306-322).
//
// With async_output on (the default, types.hpp:37) the saves go through a
// writer thread as in AsyncOutputManager (async_output.cpp): "Queued:" lines;
// with output_format 1 or 2 a record per save in output.nc (the classic
// 64-bit-offset NetCDF format, see ncwriter.cpp); with output_format 0 or 2
// ASC files in that writer's header format, Vx_%04d/Vy_%04d.asc beside
// H_%04d.asc when velocity output is on.  Each save is a device snapshot into
// one of two slots, copied to pinned host memory on a side stream while the
// steps go on, and written by the thread.
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "adaptflood.h"
#include "adaptflood_compat.hpp"

namespace {

struct Fail {
    std::string msg;
};

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail{std::string(what) + ": " + hipGetErrorString(e)};
}

void af_ok(int rc, const char* what) {
    if (rc != AF_OK) throw Fail{std::string(what) + ": " + af_last_error()};
}

template <typename T>
T* dev_alloc(size_t n) {
    void* p = nullptr;
    hip_ok(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    hip_ok(hipMemset(p, 0, (n ? n : 1) * sizeof(T)), "hipMemset");
    return static_cast<T*>(p);
}

struct Grid {
    int nx = 0, ny = 0;
    double xll = 0, yll = 0, cell = 0, nodata = 0;
};

// Stage gauges and their CSV (main.cpp:95-113, 306-322): one line per step
// that ends at t >= next_stage, "%.1f" time then ",%.4f" per gauge (H at the
// gauge cell) or ",NaN" when the cell index falls outside the grid.
struct Stages {
    FILE* fp = nullptr;
    std::vector<int64_t> idx;
    double interval = 0.0;
    void open(af_case* c, const af_par& par, int nx) {
        int32_t n = 0;
        af_ok(af_case_stage(c, -1, nullptr, nullptr, nullptr, 0, &n), "stages");
        if (!par.stage_file[0] || n == 0) return;
        const std::string path = std::string(par.output_dir) + "/stage.csv";
        fp = fopen(path.c_str(), "w");
        if (!fp) return;
        fprintf(fp, "Time(s)");
        for (int32_t i = 0; i < n; i++) {
            int32_t gx = 0, gy = 0;
            char name[256];
            af_ok(af_case_stage(c, i, &gx, &gy, name, sizeof name, &n), "stage");
            fprintf(fp, ",%s", name);
            idx.push_back((int64_t)gx + (int64_t)gy * nx);
        }
        fprintf(fp, "\n");
        interval = par.stage_int;
        printf("Stage output file: %s\n", path.c_str());
    }
    // Closes the CSV; true when one was open.
    bool close() {
        if (!fp) return false;
        fclose(fp);
        fp = nullptr;
        return true;
    }
    // writes one line from the gauge values
    void line(double t, const std::vector<double>& v) {
        fprintf(fp, "%.1f", t);
        for (double x : v) {
            if (x != x) fprintf(fp, ",NaN");
            else fprintf(fp, ",%.4f", x);
        }
        fprintf(fp, "\n");
        fflush(fp);
    }
    ~Stages() {
        if (fp) fclose(fp);
    }
};

struct Output {
    std::string dir;
    Grid g;
    bool quiet = false;
    Stages* stages = nullptr;
    void asc(const char* name, const std::vector<double>& data, const char* note = "") const {
        const std::string path = dir + "/" + name;
        af_ok(af_write_asc(path.c_str(), data.data(), g.nx, g.ny, g.xll, g.yll, g.cell),
              "write_asc");
        if (!quiet) printf("Saved: %s%s\n", path.c_str(), note);
    }
};

// The asynchronous writer.  A save takes a free slot (waiting for the
// thread when both are busy), fills its device buffers on the compute stream
// (a snapshot), and submits it: the side stream waits for the snapshot,
// copies the buffers to the slot's pinned host memory, and the thread waits
// for that copy, writes the files and frees the slot.
class AsyncWriter {
  public:
    struct Slot {
        double* dev[3] = {nullptr, nullptr, nullptr};   // H, Vx, Vy (device)
        double* host[3] = {nullptr, nullptr, nullptr};  // pinned copies
        hipEvent_t snap = nullptr, copied = nullptr;
        bool busy = false;
    };

    AsyncWriter(const Output& out, bool velocity, bool asc, bool netcdf)
        : out_(out), velocity_(velocity), asc_(asc), n_((size_t)out.g.nx * out.g.ny) {
        if (netcdf) {  // async_output.cpp:18-54
            const std::string path = out.dir + "/output.nc";
            af_ok(af_nc_create(path.c_str(), out.g.nx, out.g.ny, out.g.xll, out.g.yll,
                               out.g.cell, out.g.nodata, velocity ? 1 : 0, &nc_),
                  "nc_create");
            printf("NetCDF file created: %s\n", path.c_str());
            fflush(stdout);
        }
        hip_ok(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "hipStreamCreate");
        const int fields = velocity_ ? 3 : 1;
        for (Slot& s : slots_) {
            for (int f = 0; f < fields; f++) {
                hip_ok(hipMalloc(&s.dev[f], n_ * 8), "hipMalloc snapshot");
                hip_ok(hipHostMalloc(&s.host[f], n_ * 8, hipHostMallocDefault),
                       "hipHostMalloc snapshot");
            }
            hip_ok(hipEventCreateWithFlags(&s.snap, hipEventDisableTiming), "hipEventCreate");
            hip_ok(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming), "hipEventCreate");
        }
        thread_ = std::thread(&AsyncWriter::loop, this);
    }

    ~AsyncWriter() {
        stop();
        for (Slot& s : slots_) {
            for (int f = 0; f < 3; f++) {
                if (s.dev[f]) (void)hipFree(s.dev[f]);
                if (s.host[f]) (void)hipHostFree(s.host[f]);
            }
            if (s.snap) (void)hipEventDestroy(s.snap);
            if (s.copied) (void)hipEventDestroy(s.copied);
        }
        if (copy_) (void)hipStreamDestroy(copy_);
        if (nc_) (void)af_nc_close(nc_);
    }

    bool velocity() const { return velocity_; }

    // A free slot; waits while the thread still holds both.
    Slot& acquire() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            for (Slot& s : slots_)
                if (!s.busy) {
                    s.busy = true;
                    return s;
                }
            cv_.wait(lk);
        }
    }

    // The slot's device buffers, the state at time t, are (being) filled on
    // `compute`.
    void submit(Slot& s, int save_num, double t, hipStream_t compute) {
        hip_ok(hipEventRecord(s.snap, compute), "hipEventRecord");
        hip_ok(hipStreamWaitEvent(copy_, s.snap, 0), "hipStreamWaitEvent");
        const int fields = velocity_ ? 3 : 1;
        for (int f = 0; f < fields; f++)
            hip_ok(hipMemcpyAsync(s.host[f], s.dev[f], n_ * 8, hipMemcpyDeviceToHost, copy_),
                   "snapshot copy");
        hip_ok(hipEventRecord(s.copied, copy_), "hipEventRecord");
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back({&s, save_num, t});
        }
        cv_.notify_all();
    }

    // Waits until every submitted save is written (AsyncOutputManager::finish);
    // fails if a write failed.
    void finish() {
        stop();
        if (nc_) {
            const int rc = af_nc_close(nc_);
            nc_ = nullptr;
            if (rc != AF_OK && error_.empty()) error_ = std::string("nc_close: ") + af_last_error();
        }
        if (!error_.empty()) throw Fail{error_};
    }

  private:
    struct Job {
        Slot* slot;
        int save_num;
        double t;
    };

    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stopping_) return;
            stopping_ = true;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
    }

    void loop() {
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !jobs_.empty() || stopping_; });
                if (jobs_.empty()) return;
                job = jobs_.front();
                jobs_.pop_front();
            }
            write(job);
            {
                std::lock_guard<std::mutex> lk(mu_);
                job.slot->busy = false;
            }
            cv_.notify_all();
        }
    }

    void write(const Job& job) {
        const hipError_t e = hipEventSynchronize(job.slot->copied);
        if (e != hipSuccess) {
            note(std::string("snapshot copy: ") + hipGetErrorString(e));
            return;
        }
        if (nc_) {  // async_output.cpp:110-139
            if (!out_.quiet) {
                printf("[AsyncIO] Writing t=%.0fs to NetCDF (time_index=%d)...\n", job.t,
                       nc_records_);
                fflush(stdout);
            }
            const double* const* h = job.slot->host;
            if (af_nc_append(nc_, job.t, h[0], velocity_ ? h[1] : nullptr,
                             velocity_ ? h[2] : nullptr) != AF_OK)
                note(std::string("NetCDF: ") + af_last_error());
            else if (!out_.quiet) {
                printf("[AsyncIO] Written t=%.0fs to NetCDF, synced.\n", job.t);
                fflush(stdout);
            }
            nc_records_++;
        }
        if (!asc_) return;  // output_format 1: NetCDF only (async_output.cpp:144)
        static const char* names[3] = {"H", "Vx", "Vy"};
        const int fields = velocity_ ? 3 : 1;
        for (int f = 0; f < fields; f++) {
            char name[64];
            snprintf(name, sizeof name, "%s_%04d.asc", names[f], job.save_num);
            const std::string path = out_.dir + "/" + name;
            if (af_write_asc_async(path.c_str(), job.slot->host[f], out_.g.nx, out_.g.ny,
                                   out_.g.xll, out_.g.yll, out_.g.cell, out_.g.nodata) != AF_OK)
                note(std::string("write ") + path + ": " + af_last_error());
        }
    }

    void note(const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu_);
        if (error_.empty()) error_ = msg;
    }

    const Output& out_;
    const bool velocity_, asc_;
    const size_t n_;
    af_nc* nc_ = nullptr;
    int nc_records_ = 0;
    hipStream_t copy_ = nullptr;
    Slot slots_[2];
    std::deque<Job> jobs_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stopping_ = false;
    std::string error_;
    std::thread thread_;
};

void queued_line(const Output& out, double t, int save_num) {
    if (out.quiet) return;
    if (save_num == 0) printf("Queued: t=0s (initial state)\n");
    else printf("Queued: t=%.0fs (save_num=%d)\n", t, save_num);
    fflush(stdout);
}

void report_line(double t, double dt, double vol, double area) {
    printf("t=%.0fs: Tstep=%.2f, Vol=%.0fm3, Area=%.0fm2\n", t, dt, vol, area);
    fflush(stdout);
}

// Host volume and flooded area of a host copy of H (main.cpp:292-298).
void mass_of(const std::vector<double>& H, double dA, double* vol, double* area) {
    double v = 0.0, a = 0.0;
    for (double h : H) {
        v += h * dA;
        if (h > 0.01) a += dA;
    }
    *vol = v;
    *area = a;
}

// Device velocity from a reference-layout state (kernels.cpp:293-351).
void velocity_files(const Output& out, const std::vector<double>& H, const std::vector<double>& dem,
                    const std::vector<double>& Qx, const std::vector<double>& Qy, double dx) {
    const size_t n = H.size(), nq = Qx.size();
    double* dH = dev_alloc<double>(n);
    double* dZ = dev_alloc<double>(n);
    double* dQx = dev_alloc<double>(nq);
    double* dQy = dev_alloc<double>(nq);
    double* dVx = dev_alloc<double>(n);
    double* dVy = dev_alloc<double>(n);
    hip_ok(hipMemcpy(dH, H.data(), n * 8, hipMemcpyHostToDevice), "copy H");
    hip_ok(hipMemcpy(dZ, dem.data(), n * 8, hipMemcpyHostToDevice), "copy DEM");
    hip_ok(hipMemcpy(dQx, Qx.data(), nq * 8, hipMemcpyHostToDevice), "copy Qx");
    hip_ok(hipMemcpy(dQy, Qy.data(), nq * 8, hipMemcpyHostToDevice), "copy Qy");
    ComputeVelocity_GPU(dVx, dVy, dH, dZ, dQx, dQy, out.g.nx, out.g.ny, dx, out.g.nodata);
    std::vector<double> vx(n), vy(n);
    hip_ok(hipMemcpy(vx.data(), dVx, n * 8, hipMemcpyDeviceToHost), "copy Vx");
    hip_ok(hipMemcpy(vy.data(), dVy, n * 8, hipMemcpyDeviceToHost), "copy Vy");
    for (void* p : {(void*)dH, (void*)dZ, (void*)dQx, (void*)dQy, (void*)dVx, (void*)dVy})
        (void)hipFree(p);
    out.asc("Vx_final.asc", vx);
    out.asc("Vy_final.asc", vy);
}

struct Result {
    double t = 0.0;
    long long steps = 0;
    std::vector<double> H, Qx, Qy;
};

// The reference's time loop through the drop-in wrappers (main.cpp:196-324).
Result run_compat(const af_scenario_desc& d, const af_par& par, const Output& out,
                  AsyncWriter* aw) {
    const int nx = d.nx, ny = d.ny;
    const size_t n = (size_t)nx * ny, nq = (size_t)(nx + 1) * (ny + 1);
    const double dx = d.dx, dA = dx * dx;
    double* H = dev_alloc<double>(n);
    double* DEM = dev_alloc<double>(n);
    double* Qx = dev_alloc<double>(nq);
    double* Qy = dev_alloc<double>(nq);
    double* Qxold = dev_alloc<double>(nq);
    double* Qyold = dev_alloc<double>(nq);
    void* ws = dev_alloc<char>(AF_FLOODPLAINQ_WORKSPACE_BYTES);
    hip_ok(hipMemcpy(DEM, d.dem, n * 8, hipMemcpyHostToDevice), "copy DEM");
    if (d.h0) hip_ok(hipMemcpy(H, d.h0, n * 8, hipMemcpyHostToDevice), "copy H0");
    af::compat::bind(nullptr, ws);

    // flow (QFIX/QVAR) and other (FREE/HFIX/HVAR) points, main.cpp:77-85
    std::vector<int> flow, other;
    for (int p = 0; p < d.n_points; p++) {
        const int ty = d.pt_type[p];
        if (ty == AF_BC_QFIX || ty == AF_BC_QVAR) flow.push_back(p);
        else other.push_back(p);
    }
    auto idx_of = [&](int p) { return d.pt_x[p] + d.pt_y[p] * nx; };
    // the per-function kernels index H with these directly (as main.cpp:
    // 253-259 and kernels.cpp:263-272 do); a point outside the grid is an error
    for (int p = 0; p < d.n_points; p++) {
        const long long k = (long long)d.pt_x[p] + (long long)d.pt_y[p] * nx;
        if (k < 0 || k >= (long long)nx * ny)
            throw Fail{"point " + std::to_string(p) + " at grid (" + std::to_string(d.pt_x[p]) +
                       ", " + std::to_string(d.pt_y[p]) + ") lies outside the grid"};
    }
    auto value_at = [&](int p, double t) {
        const int len = d.pt_series_len ? d.pt_series_len[p] : 0;
        const int off = d.pt_series_off ? d.pt_series_off[p] : 0;
        return af_interp_series(d.series_t + off, d.series_v + off, len, d.pt_value[p], t);
    };
    const int nf = (int)flow.size(), no = (int)other.size();
    std::vector<int32_t> ps_idx_h(nf), bc_idx_h(no), bc_type_h(no);
    std::vector<double> ps_flow_h(nf), bc_value_h(no);
    for (int s = 0; s < nf; s++) ps_idx_h[s] = idx_of(flow[s]);
    for (int b = 0; b < no; b++) {
        const int p = other[b], ty = d.pt_type[p];
        bc_idx_h[b] = idx_of(p);
        bc_type_h[b] = ty == AF_BC_FREE ? 1 : ty == AF_BC_HFIX ? 2 : ty == AF_BC_HVAR ? 3 : 0;
        bc_value_h[b] = d.pt_value[p];
    }
    int32_t* ps_idx = dev_alloc<int32_t>(nf);
    double* ps_flow = dev_alloc<double>(nf);
    int32_t* bc_idx = dev_alloc<int32_t>(no);
    int32_t* bc_type = dev_alloc<int32_t>(no);
    double* bc_value = dev_alloc<double>(no);
    if (nf) hip_ok(hipMemcpy(ps_idx, ps_idx_h.data(), nf * 4, hipMemcpyHostToDevice), "ps_idx");
    if (no) {
        hip_ok(hipMemcpy(bc_idx, bc_idx_h.data(), no * 4, hipMemcpyHostToDevice), "bc_idx");
        hip_ok(hipMemcpy(bc_type, bc_type_h.data(), no * 4, hipMemcpyHostToDevice), "bc_type");
    }

    std::vector<double> Hh(n);
    auto pull_H = [&]() { hip_ok(hipMemcpy(Hh.data(), H, n * 8, hipMemcpyDeviceToHost), "H"); };
    int save_num = 0;
    double t = 0.0;
    // main.cpp:206-225 (t = 0) and 326-363
    auto save = [&]() {
        if (aw) {
            AsyncWriter::Slot& sl = aw->acquire();
            hip_ok(hipMemcpyAsync(sl.dev[0], H, n * 8, hipMemcpyDeviceToDevice, nullptr),
                   "snapshot H");
            if (aw->velocity())
                ComputeVelocity_GPU(sl.dev[1], sl.dev[2], H, DEM, Qx, Qy, nx, ny, dx, d.nodata);
            aw->submit(sl, save_num, t, nullptr);
            queued_line(out, t, save_num);
        } else {
            char name[64];
            pull_H();
            snprintf(name, sizeof name, "H_%04d.asc", save_num);
            out.asc(name, Hh, save_num == 0 ? " (initial state)" : "");
        }
        save_num++;
    };
    save();

    double next_save = par.save_int, next_mass = par.mass_int;
    double next_stage = par.stage_int;
    long long step = 0;
    while (t < d.sim_time) {
        double Tstep = d.init_tstep;
        FloodplainQ_GPU(H, DEM, Qx, Qy, Qxold, Qyold, nx, ny, dx, dx, d.fpn, d.depth_thresh,
                        d.max_hflow, 0.01, Tstep, 1.0, dA, d.cfl, d.g, d.init_tstep, d.theta,
                        d.nodata);
        BCs_GPU(Qx, Qy, H, DEM, nx, ny, dx, d.g, Tstep, d.nodata, d.west_free, d.east_free,
                d.north_free, d.south_free);
        if (!out.quiet && (step < 10 || step % 100 == 0)) {
            printf("Step %lld, t=%.2fs, dt=%.4fs\n", step, t, Tstep);
            fflush(stdout);
        }
        if (nf) {
            for (int s = 0; s < nf; s++) ps_flow_h[s] = value_at(flow[s], t) * dx * Tstep / dA;
            hip_ok(hipMemcpy(ps_flow, ps_flow_h.data(), nf * 8, hipMemcpyHostToDevice), "ps_flow");
            AddFlowBCs_GPU(H, ps_idx, ps_flow, nf);
        }
        if (d.rain_enabled) {
            const double r = af_interp_series(d.rain_t, d.rain_r, d.n_rain, 0.0, t);
            Rainfall_GPU(H, DEM, nx, ny, r, Tstep, d.nodata);
        }
        UpdateH_GPU(H, Qx, Qy, nx, ny, dA, Tstep);
        if (no) {
            for (int b = 0; b < no; b++)
                if (d.pt_type[other[b]] == AF_BC_HVAR) bc_value_h[b] = value_at(other[b], t);
            hip_ok(hipMemcpy(bc_value, bc_value_h.data(), no * 8, hipMemcpyHostToDevice),
                   "bc_value");
            ApplyPointBCs_GPU(H, bc_idx, bc_type, bc_value, no, nx, dx, dA, Tstep);
        }
        UpdateQs_GPU(Qx, Qy, Qxold, Qyold, nx, ny, dx);
        t += Tstep;
        step++;
        if (t >= next_mass) {
            pull_H();
            double vol, area;
            mass_of(Hh, dA, &vol, &area);
            report_line(t, Tstep, vol, area);
            next_mass += d.mass_int;
        }
        if (out.stages && out.stages->fp && t >= next_stage) {
            std::vector<double> v(out.stages->idx.size());
            const long long ncell = (long long)nx * ny;
            for (size_t g = 0; g < v.size(); g++) {
                const int64_t k = out.stages->idx[g];
                v[g] = NAN;
                if (k >= 0 && k < ncell)
                    hip_ok(hipMemcpy(&v[g], H + k, 8, hipMemcpyDeviceToHost), "gauge");
            }
            out.stages->line(t, v);
            next_stage += out.stages->interval;
        }
        if (t >= next_save) {
            save();
            next_save += par.save_int;
        }
    }
    Result r;
    r.t = t;
    r.steps = step;
    pull_H();
    r.H = Hh;
    r.Qx.resize(nq);
    r.Qy.resize(nq);
    hip_ok(hipMemcpy(r.Qx.data(), Qx, nq * 8, hipMemcpyDeviceToHost), "Qx");
    hip_ok(hipMemcpy(r.Qy.data(), Qy, nq * 8, hipMemcpyDeviceToHost), "Qy");
    for (void* p : {(void*)H, (void*)DEM, (void*)Qx, (void*)Qy, (void*)Qxold, (void*)Qyold, ws,
                    (void*)ps_idx, (void*)ps_flow, (void*)bc_idx, (void*)bc_type, (void*)bc_value})
        (void)hipFree(p);
    return r;
}

// The fused domain: one launch per step, graph-replayed, stopping at each
// save time.  Mass reports are recorded on the device at the exact steps.
Result run_fused(const af_scenario_desc& d, const af_par& par, const Output& out, int precision,
                 AsyncWriter* aw) {
    const size_t n = (size_t)d.nx * d.ny, nq = (size_t)(d.nx + 1) * (d.ny + 1);
    af_tuning tun{};
    // one record per report; massint <= 0 reports after every step
    // (main.cpp:291-303), a count not known in advance: then the largest
    // buffer (1M records, 32 MiB), and reports past it are counted, not shown
    const double nrep = d.mass_int > 0 ? std::ceil(d.sim_time / d.mass_int) + 4 : 1048576.0;
    tun.report_cap = (int32_t)std::min(nrep, 1048576.0);
    // dt of the first 4M steps, for the Step lines
    const bool step_lines = !out.quiet;
    tun.dt_log_cap = step_lines ? (1 << 22) : 0;
    size_t bytes = 0;
    af_ok(af_domain_arena_bytes(&d, precision, &tun, &bytes), "arena_bytes");
    void* arena = nullptr;
    hip_ok(hipMalloc(&arena, bytes), "hipMalloc arena");
    af_domain* D = nullptr;
    af_ok(af_domain_create(&d, precision, &tun, arena, bytes, nullptr, &D), "domain_create");

    std::vector<double> Hh(n), rep;
    int save_num = 0, shown = 0;
    af_run_info info;
    af_ok(af_domain_info(D, &info, nullptr), "info");
    // main.cpp:206-225 (t = 0) and 326-363
    auto save = [&]() {
        if (aw) {
            AsyncWriter::Slot& sl = aw->acquire();
            const bool v = aw->velocity();
            af_ok(af_batch_snapshot(D, 0, sl.dev[0], v ? sl.dev[1] : nullptr,
                                    v ? sl.dev[2] : nullptr, nullptr),
                  "snapshot");
            aw->submit(sl, save_num, info.t, nullptr);
            queued_line(out, info.t, save_num);
        } else {
            char name[64];
            af_ok(af_domain_read(D, 0, Hh.data(), nullptr), "read H");
            snprintf(name, sizeof name, "H_%04d.asc", save_num);
            out.asc(name, Hh, save_num == 0 ? " (initial state)" : "");
        }
        save_num++;
    };
    save();

    // The "Step" lines of main.cpp:238-241 come from the device dt log, in
    // step order with the mass-report lines: the host sums the logged dt in
    // sequence, as the device advances t, so the times agree exactly and a
    // report (recorded at the t that ends its step) follows its step's line.
    std::vector<double> dts;
    long long printed = 0;  // steps whose lines are out
    double t_run = 0.0;     // start time of step `printed`
    bool dropped_line = false;
    auto show_progress = [&]() {
        af_run_info now;
        af_ok(af_domain_info(D, &now, nullptr), "info");
        // records held on the device: the first report_cap of them
        const int32_t held = std::min(now.n_reports, tun.report_cap);
        int32_t nrep = shown;
        if (held > shown) {
            rep.assign(4 * (size_t)held, 0.0);
            int32_t total = 0;
            af_ok(af_domain_read_reports(D, rep.data(), held, &total, nullptr), "reports");
            nrep = held;
        }
        int64_t logged = 0;
        if (step_lines) {
            af_ok(af_domain_read_dt_log(D, nullptr, 0, &logged, nullptr), "dt log");
            dts.resize((size_t)logged);
            if (logged > printed)
                af_ok(af_domain_read_dt_log(D, dts.data(), logged, &logged, nullptr), "dt log");
        }
        for (; printed < logged; printed++) {
            const double dt = dts[(size_t)printed];
            if (printed < 10 || printed % 100 == 0)
                printf("Step %lld, t=%.2fs, dt=%.4fs\n", printed, t_run, dt);
            t_run += dt;
            while (shown < nrep && rep[4 * (size_t)shown] == t_run) {
                report_line(rep[4 * shown], rep[4 * shown + 1], rep[4 * shown + 2],
                            rep[4 * shown + 3]);
                shown++;
            }
        }
        for (; shown < nrep; shown++)  // past the logged steps
            report_line(rep[4 * shown], rep[4 * shown + 1], rep[4 * shown + 2],
                        rep[4 * shown + 3]);
        if (now.n_reports > held && !dropped_line) {
            printf("(mass reports after the first %d are not recorded)\n", held);
            dropped_line = true;
        }
        fflush(stdout);
    };

    double next_save = par.save_int;
    const bool gauges = out.stages && out.stages->fp;
    double next_stage = par.stage_int;
    while (info.t < d.sim_time) {
        double stop = d.sim_time;
        if (par.save_int > 0 && next_save < stop) stop = next_save;
        // saveint <= 0 saves after every step, as main.cpp:326 does when
        // next_save never passes t
        if (par.save_int <= 0) stop = info.t;
        if (gauges && next_stage < stop) stop = next_stage;  // interval 0: every step
        if (info.t >= stop) {
            // one step jumped past two output times: the reference writes
            // again after the next step
            af_ok(af_batch_set_sim_time(D, 0, d.sim_time, nullptr), "set_sim_time");
            af_ok(af_domain_advance(D, 1, 0, nullptr), "advance");
        } else {
            af_ok(af_batch_set_sim_time(D, 0, stop, nullptr), "set_sim_time");
            af_ok(af_domain_run(D, 216, 1, nullptr), "run");
        }
        af_ok(af_domain_info(D, &info, nullptr), "info");
        show_progress();
        if (gauges && info.t >= next_stage) {
            std::vector<double> v(out.stages->idx.size());
            af_ok(af_batch_read_cells(D, 0, out.stages->idx.data(), (int)v.size(), v.data(),
                                      nullptr),
                  "gauges");
            out.stages->line(info.t, v);
            next_stage += out.stages->interval;
        }
        if (info.t >= next_save) {
            save();
            next_save += par.save_int;
        }
    }
    Result r;
    r.t = info.t;
    r.steps = info.steps_done;
    r.H.resize(n);
    r.Qx.resize(nq);
    r.Qy.resize(nq);
    af_ok(af_domain_read(D, 0, r.H.data(), nullptr), "read H");
    af_ok(af_domain_read(D, 1, r.Qx.data(), nullptr), "read Qx");
    af_ok(af_domain_read(D, 2, r.Qy.data(), nullptr), "read Qy");
    af_domain_destroy(D);
    (void)hipFree(arena);
    return r;
}

int usage(const char* argv0) {
    fprintf(stderr, "usage: %s [--path compat|fp64|fp32] [--quiet] [parfile]\n", argv0);
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    std::string path = "compat";
    const char* parfile = nullptr;
    bool quiet = false;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--path") && i + 1 < argc) path = argv[++i];
        else if (!strcmp(argv[i], "--quiet")) quiet = true;
        else if (argv[i][0] == '-') return usage(argv[0]);
        else parfile = argv[i];
    }
    if (path != "compat" && path != "fp64" && path != "fp32") return usage(argv[0]);
    af_case* c = af_case_new();
    try {
        // console lines of main.cpp:19-46, 70-86
        printf("=== LISFLOOD-GPU Flood Simulation ===\n");
        printf("Execution space: HIP gfx950 (%s path)\n\n", path.c_str());
        af_par par;
        if (parfile) {
            printf("Reading parameter file: %s\n", parfile);
            af_ok(af_read_par(parfile, &par), "load");
        } else {  // main.cpp:28-31: the defaults, inputs from the working directory
            printf("Usage: %s <parfile>\n", argv[0]);
            printf("Using default parameters...\n\n");
            af_par_defaults(&par);
        }
        static const char* fmt_str[] = {"ASC", "NetCDF", "Both"};
        printf("Parameters:\n");
        printf("  DEM file:    %s\n", par.dem_file);
        printf("  BCI file:    %s\n", par.bci_file);
        printf("  BDY file:    %s\n", par.bdy_file);
        printf("  Output dir:  %s\n", par.output_dir);
        printf("  Sim time:    %.0f s\n", par.sim_time);
        printf("  Save int:    %.0f s\n", par.save_int);
        printf("  CFL:         %.2f (adaptive timestep)\n", par.cfl);
        printf("  Manning n:   %.4f\n", par.fpn);
        printf("  Output fmt:  %s\n",
               par.output_format >= 0 && par.output_format <= 2 ? fmt_str[par.output_format]
                                                                 : "?");
        printf("  Velocity:    %s\n", par.voutput ? "enabled" : "disabled");
        printf("  Async I/O:   %s\n", par.async_output ? "enabled" : "disabled");
        printf("\n");
        fflush(stdout);
        af_ok(af_case_load_par(c, &par), "load");
        af_ok(af_case_par(c, &par), "par");
        af_scenario_desc d;
        af_ok(af_case_scenario(c, &d), "scenario");
        if (d.west_free || d.east_free || d.north_free || d.south_free)
            printf("Line FREE boundaries: W=%d E=%d N=%d S=%d\n", d.west_free ? 1 : 0,
                   d.east_free ? 1 : 0, d.north_free ? 1 : 0, d.south_free ? 1 : 0);
        size_t n_flow = 0;
        for (int p = 0; p < d.n_points; p++)
            if (d.pt_type[p] == AF_BC_QFIX || d.pt_type[p] == AF_BC_QVAR) n_flow++;
        printf("Flow boundaries: %zu, Other boundaries: %zu\n", n_flow,
               (size_t)d.n_points - n_flow);
        Output out;
        out.quiet = quiet;
        out.dir = par.output_dir;
        af_ok(af_case_dem_header(c, &out.g.nx, &out.g.ny, &out.g.xll, &out.g.yll, &out.g.cell,
                                 &out.g.nodata),
              "dem header");
        mkdir(par.output_dir, 0755);
        Stages stages;
        stages.open(c, par, d.nx);
        out.stages = &stages;
        printf("\nStarting simulation: %.0f seconds, output every %.0f seconds\n\n", d.sim_time,
               par.save_int);
        fflush(stdout);
        if (par.output_format >= 1 && !par.async_output)
            printf("Note: output_format %d: NetCDF is written by the asynchronous writer only "
                   "(async_output 1), as in the reference; none is written\n",
                   par.output_format);
        const auto w0 = std::chrono::steady_clock::now();
        std::unique_ptr<AsyncWriter> aw;
        if (par.async_output)
            aw.reset(new AsyncWriter(out, par.voutput != 0, par.output_format != 1,
                                     par.output_format >= 1));
        Result r = path == "compat"
                       ? run_compat(d, par, out, aw.get())
                       : run_fused(d, par, out, path == "fp64" ? AF_FP64 : AF_FP32, aw.get());
        const double wall =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
        if (aw) {
            printf("\nWaiting for async output to complete...\n");
            fflush(stdout);
            aw->finish();
            printf("Async output finished.\n");
        }
        out.asc("H_final.asc", r.H);
        if (par.voutput) {
            std::vector<double> dem(d.dem, d.dem + (size_t)d.nx * d.ny);
            velocity_files(out, r.H, dem, r.Qx, r.Qy, d.dx);
        }
        if (stages.close()) printf("Stage output saved.\n");  // main.cpp:390-394
        double vol, area;
        mass_of(r.H, d.dx * d.dx, &vol, &area);
        printf("\n=== Simulation Complete ===\n");
        printf("Simulation time: %.0f s, Steps: %lld\n", r.t, r.steps);
        printf("Wall clock time: %.2f s (%.2f min)\n", wall, wall / 60.0);
        printf("Final volume: %.0f m3\n", vol);
        printf("Output saved to %s/\n", par.output_dir);
        if (par.output_format >= 1 && par.async_output)
            printf("NetCDF output: %s/output.nc\n", par.output_dir);
    } catch (const Fail& f) {
        fprintf(stderr, "lisflood_amd: %s\n", f.msg.c_str());
        af_case_free(c);
        return 1;
    } catch (const af::compat::error& e) {
        fprintf(stderr, "lisflood_amd: %s\n", e.what());
        af_case_free(c);
        return 1;
    }
    af_case_free(c);
    return 0;
}
