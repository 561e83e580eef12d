// compat.hip -- the reference call surface (src/kernels.hpp:8-37 and the
// inline AddFlowBCs of src/main.cpp:256-260) as FP64 HIP kernels on the
// reference memory layout.  These exist so a caller of the reference's
// functions can switch over one function at a time; the fused step in
// fused.hip is the production path and shares the arithmetic in
// af_device.hpp.
#include <float.h>

#include <algorithm>
#include <atomic>

#include "af_device.hpp"
#include "af_host.hpp"

namespace af {
namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads)
k_update_qs(const double* __restrict__ Qx, const double* __restrict__ Qy,
            double* __restrict__ Qxold, double* __restrict__ Qyold, long long n, double dxinv) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        Qxold[k] = Qx[k] * dxinv;
        Qyold[k] = Qy[k] * dxinv;
    }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// H max over cells with DEM != nodata (kernels.cpp:30-33); NaN never wins
// because `H > m` is false for NaN, as in the reference.
__global__ void __launch_bounds__(kThreads)
k_hmax(const double* __restrict__ H, const double* __restrict__ DEM, long long n, double nodata,
       unsigned long long* __restrict__ out) {
    double m = -DBL_MAX;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        double h = H[k];
        if (DEM[k] != nodata && h > m) m = h;
    }
    for (int off = 32; off > 0; off >>= 1) m = getmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, fkey(m));
}

// x faces (kernels.cpp:46-93): i in [0, xsz-2], j in [0, ysz-1]
__global__ void __launch_bounds__(kThreads)
k_qx(const double* __restrict__ H, const double* __restrict__ DEM, double* __restrict__ Qx,
     const double* __restrict__ Qxold, const double* __restrict__ Qyold, int xsz, int ysz,
     double nodata, FluxConstsF64 k) {
    const long long n = (long long)(xsz - 1) * ysz;
    const long long P = xsz + 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % (xsz - 1));
        long long j = t / (xsz - 1);
        long long p0 = i + j * xsz, p1 = p0 + 1;
        long long pq0 = (i + 1) + j * P;
        double z0 = DEM[p0], z1 = DEM[p1];
        if (z0 == nodata || z1 == nodata) {
            Qx[pq0] = 0.0;
            continue;
        }
        double q0 = Qxold[pq0];
        double qup = (i > 0) ? Qxold[pq0 - 1] : 0.0;
        double qdown = (i < xsz - 2) ? Qxold[pq0 + 1] : 0.0;
        double qt = (Qyold[i + j * P] + Qyold[(i + 1) + j * P] + Qyold[i + (j + 1) * P] +
                     Qyold[(i + 1) + (j + 1) * P]) / 4.0;
        Qx[pq0] = face_flux_ref(z0, z1, H[p0], H[p1], q0, qup, qdown, qt, k);
    }
}

// y faces (kernels.cpp:96-143): i in [0, xsz-1], j in [0, ysz-2]
__global__ void __launch_bounds__(kThreads)
k_qy(const double* __restrict__ H, const double* __restrict__ DEM, double* __restrict__ Qy,
     const double* __restrict__ Qxold, const double* __restrict__ Qyold, int xsz, int ysz,
     double nodata, FluxConstsF64 k) {
    const long long n = (long long)xsz * (ysz - 1);
    const long long P = xsz + 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % xsz);
        long long j = t / xsz;
        long long p0 = i + j * xsz, p1 = i + (j + 1) * xsz;
        long long pq0 = i + (j + 1) * P;
        double z0 = DEM[p0], z1 = DEM[p1];
        if (z0 == nodata || z1 == nodata) {
            Qy[pq0] = 0.0;
            continue;
        }
        double q0 = Qyold[pq0];
        double qup = (j > 0) ? Qyold[i + j * P] : 0.0;
        double qdown = (j < ysz - 2) ? Qyold[i + (j + 2) * P] : 0.0;
        double qt = (Qxold[i + j * P] + Qxold[(i + 1) + j * P] + Qxold[i + (j + 1) * P] +
                     Qxold[(i + 1) + (j + 1) * P]) / 4.0;
        Qy[pq0] = face_flux_ref(z0, z1, H[p0], H[p1], q0, qup, qdown, qt, k);
    }
}

// All four edges of BCs_GPU (kernels.cpp:149-256) in one launch.  A FREE side
// writes its xsz or ysz edge faces; a CLOSED side zeroes xsz+1 or ysz+1
// entries, including the padding entry, exactly as the reference does.
__global__ void __launch_bounds__(kThreads)
k_bcs(double* __restrict__ Qx, double* __restrict__ Qy, const double* __restrict__ H,
      const double* __restrict__ DEM, int xsz, int ysz, double dx, double g, double nodata,
      int wf, int ef, int nf, int sf) {
    const long long P = xsz + 1;
    const long long nw = wf ? ysz : ysz + 1, ne = ef ? ysz : ysz + 1;
    const long long nn = nf ? xsz : xsz + 1, ns = sf ? xsz : xsz + 1;
    const long long total = nw + ne + nn + ns;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long u = t;
        if (u < nw) {  // west, face j*P
            long long j = u;
            double q = 0.0;
            if (wf) {
                long long p0 = j * xsz, p1 = 1 + j * xsz;
                double z0 = DEM[p0];
                q = edge_west<double>(z0, H[p0], DEM[p1], H[p1], z0 == nodata, dx, g);
            }
            Qx[j * P] = q;
            continue;
        }
        u -= nw;
        if (u < ne) {  // east, face xsz + j*P
            long long j = u;
            double q = 0.0;
            if (ef) {
                long long p0 = (xsz - 1) + j * xsz;
                double z0 = DEM[p0];
                q = edge_out_pos<double>(H[p0], z0 == nodata, dx, g);
            }
            Qx[xsz + j * P] = q;
            continue;
        }
        u -= ne;
        if (u < nn) {  // north, face i
            long long i = u;
            double q = 0.0;
            if (nf) q = edge_out_neg<double>(H[i], DEM[i] == nodata, dx, g);
            Qy[i] = q;
            continue;
        }
        u -= nn;
        {  // south, face i + ysz*P
            long long i = u;
            double q = 0.0;
            if (sf) {
                long long p0 = i + (long long)(ysz - 1) * xsz;
                q = edge_out_pos<double>(H[p0], DEM[p0] == nodata, dx, g);
            }
            Qy[i + (long long)ysz * P] = q;
        }
    }
}

// Point lists are grouped by cell with one sort in LDS per chunk of up to
// kSortCap entries: key = cell << 32 | position in the chunk, so a stable
// order within a cell is the array order.  The first entry of each run of
// equal cells then applies the whole run, in array order.  Longer lists are
// taken chunk after chunk in array order on one stream, so every cell still
// sees its entries in array order: O(n log^2 kSortCap) work in all, where a
// scan for each entry's duplicates would be O(n^2).
constexpr int kSortCap = 8192;       // 64 KiB of keys in LDS
constexpr int kSortThreads = 1024;
constexpr unsigned long long kNoKey = ~0ull;

// Sorts key[0, m) ascending (m a power of two) with a bitonic network; every
// thread of the block takes part and the block is synchronised on return.
__device__ void lds_bitonic_sort(unsigned long long* key, int m) {
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = key[i], b = key[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        key[i] = b;
                        key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int pow2_at_least(int n) {
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

// AddFlowBCs (main.cpp:256-260) for one chunk of n <= kSortCap sources.
__global__ void __launch_bounds__(kSortThreads)
k_add_flow_sorted(double* __restrict__ H, const int32_t* __restrict__ idx,
                  const double* __restrict__ ps, int n) {
    __shared__ unsigned long long key[kSortCap];
    const int m = pow2_at_least(n);
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        const int c = (k < n) ? idx[k] : -1;
        key[k] = (c >= 0) ? ((unsigned long long)c << 32) | (unsigned)k : kNoKey;
    }
    __syncthreads();
    lds_bitonic_sort(key, m);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const unsigned long long e = key[k];
        if (e == kNoKey) continue;
        const unsigned c = (unsigned)(e >> 32);
        if (k > 0 && (unsigned)(key[k - 1] >> 32) == c) continue;  // not the run's first
        double h = H[c];
        for (int r = k; r < n && key[r] != kNoKey && (unsigned)(key[r] >> 32) == c; r++) {
            h += ps[(unsigned)key[r]];
            if (h < 0.0) h = 0.0;
        }
        H[c] = h;
    }
}

// ApplyPointBCs (kernels.cpp:261-274) for one chunk of n <= kSortCap points:
// the last FREE/HFIX/HVAR point of each cell in array order wins.
__global__ void __launch_bounds__(kSortThreads)
k_apply_point_bcs_sorted(double* __restrict__ H, const int32_t* __restrict__ idx,
                         const int32_t* __restrict__ type, const double* __restrict__ value,
                         int n) {
    __shared__ unsigned long long key[kSortCap];
    const int m = pow2_at_least(n);
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        int c = -1;
        if (k < n) {
            const int ty = type[k];
            if (ty == 1 || ty == 2 || ty == 3) c = idx[k];
        }
        key[k] = (c >= 0) ? ((unsigned long long)c << 32) | (unsigned)k : kNoKey;
    }
    __syncthreads();
    lds_bitonic_sort(key, m);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const unsigned long long e = key[k];
        if (e == kNoKey) continue;
        const unsigned c = (unsigned)(e >> 32);
        if (k + 1 < n && key[k + 1] != kNoKey && (unsigned)(key[k + 1] >> 32) == c) continue;
        const int b = (int)(unsigned)e;   // the run's last entry
        H[c] = (type[b] == 1) ? 0.0 : value[b];
    }
}

// UpdateH (kernels.cpp:281-288)
__global__ void __launch_bounds__(kThreads)
k_update_h(double* __restrict__ H, const double* __restrict__ Qx, const double* __restrict__ Qy,
           int xsz, int ysz, double dA, double Tstep) {
    const long long n = (long long)xsz * ysz;
    const long long P = xsz + 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long i = t % xsz, j = t / xsz;
        double Qx_in = Qx[i + j * P], Qx_out = Qx[(i + 1) + j * P];
        double Qy_in = Qy[i + j * P], Qy_out = Qy[i + (j + 1) * P];
        double h = H[t];
        h += Tstep * (Qx_in - Qx_out + Qy_in - Qy_out) / dA;
        if (h < 0.0) h = 0.0;
        H[t] = h;
    }
}

// Rainfall (kernels.cpp:360-365)
__global__ void __launch_bounds__(kThreads)
k_rainfall(double* __restrict__ H, const double* __restrict__ DEM, long long n, double depth,
           double nodata) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        if (DEM[t] != nodata) {
            double h = H[t] + depth;
            if (h < 0.0) h = 0.0;
            H[t] = h;
        }
    }
}

// ComputeVelocity (kernels.cpp:296-348)
__global__ void __launch_bounds__(kThreads)
k_velocity(double* __restrict__ Vx, double* __restrict__ Vy, const double* __restrict__ H,
           const double* __restrict__ DEM, const double* __restrict__ Qx,
           const double* __restrict__ Qy, int xsz, int ysz, double dx, double nodata) {
    const long long n = (long long)xsz * ysz;
    const long long P = xsz + 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long i = t % xsz, j = t / xsz;
        double z = DEM[t], h = H[t];
        if (z == nodata || h < 0.001) {
            Vx[t] = 0.0;
            Vy[t] = 0.0;
            continue;
        }
        double hflow = h;
        if (hflow < 0.001) hflow = 0.001;
        double a = Qx[i + j * P], b = Qx[(i + 1) + j * P];
        double va = (fabs(a) > 1e-10) ? a / dx / hflow : 0.0;
        double vb = (fabs(b) > 1e-10) ? b / dx / hflow : 0.0;
        Vx[t] = (fabs(va) > fabs(vb)) ? va : vb;
        double c = Qy[i + j * P], d = Qy[i + (j + 1) * P];
        double vc = (fabs(c) > 1e-10) ? c / dx / hflow : 0.0;
        double vd = (fabs(d) > 1e-10) ? d / dx / hflow : 0.0;
        Vy[t] = (fabs(vc) > fabs(vd)) ? vc : vd;
    }
}

int check_grid(int xsz, int ysz) {
    AF_REQUIRE(xsz >= 2 && ysz >= 2, AF_E_INVALID, "grid %d x %d: need at least 2 x 2", xsz, ysz);
    return AF_OK;
}

// Scratch of af_FloodplainQ_GPU_noalloc: the H-max key of one call (one
// 64-bit word, padded to the workspace size of af_FloodplainQ_GPU).  Each GPU
// the library runs on holds its own copy until the process ends.  Host code
// never passes `scratch` itself to a kernel (on the host it names a shadow
// variable, not device memory): it passes the device address that
// hipGetSymbolAddress returned for the current device.
constexpr int kScratchBytes = AF_FLOODPLAINQ_WORKSPACE_BYTES;
static __device__ __align__(256) unsigned long long scratch[kScratchBytes / 8];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the device `stream` runs on (the current
// device for the null stream), looked up once per device with that device
// current.
int scratch_ptr(hipStream_t stream, unsigned long long** out) {
    int cur = 0, dev = 0;
    AF_HIP(hipGetDevice(&cur));
    dev = cur;
    if (stream) AF_HIP(hipStreamGetDevice(stream, &dev));
    AF_REQUIRE(dev >= 0 && dev < kMaxDevices, AF_E_INVALID, "device %d out of range", dev);
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        if (dev != cur) AF_HIP(hipSetDevice(dev));
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur) AF_HIP(hipSetDevice(cur));
        if (e != hipSuccess) {
            set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return AF_E_HIP;
        }
        AF_REQUIRE(p != nullptr, AF_E_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *out = reinterpret_cast<unsigned long long*>(p);
    return AF_OK;
}

// FloodplainQ_GPU with the H-max key at `key` (kernels.cpp:22-146).
int floodplain_q(const double* H, const double* DEM, double* Qx, double* Qy, const double* Qxold,
                 const double* Qyold, int xsz, int ysz, double dx, double FPn, double DepthThresh,
                 double MaxHflow, double* Tstep, double cfl, double g, double InitTstep,
                 double theta, double nodata, unsigned long long* key, hipStream_t s) {
    long long n = (long long)xsz * ysz;
    // the key is cleared by a kernel, not a memset, so the call sequence
    // holds only kernel launches and one read-back
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, key, fkey(-DBL_MAX));
    unsigned hb = blocks_for(n, kThreads);
    if (hb > 4096) hb = 4096;
    hipLaunchKernelGGL(k_hmax, dim3(hb), dim3(kThreads), 0, s, H, DEM, n, nodata, key);
    AF_HIP(hipGetLastError());
    unsigned long long hkey = 0;
    AF_HIP(hipMemcpyAsync(&hkey, key, sizeof(hkey), hipMemcpyDeviceToHost, s));
    AF_HIP(hipStreamSynchronize(s));
    double Hmax = fkey_inv_d(hkey);
    double dt = cfl_dt(Hmax, DepthThresh, cfl, dx, g, InitTstep);
    *Tstep = dt;
    FluxConstsF64 k{dx, FPn, DepthThresh, MaxHflow, g * dt, theta};
    long long nx = (long long)(xsz - 1) * ysz, ny = (long long)xsz * (ysz - 1);
    hipLaunchKernelGGL(k_qx, dim3(blocks_for(nx, kThreads)), dim3(kThreads), 0, s, H, DEM, Qx,
                       Qxold, Qyold, xsz, ysz, nodata, k);
    hipLaunchKernelGGL(k_qy, dim3(blocks_for(ny, kThreads)), dim3(kThreads), 0, s, H, DEM, Qy,
                       Qxold, Qyold, xsz, ysz, nodata, k);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

}  // namespace
}  // namespace af

using namespace af;

extern "C" {

int af_UpdateQs_GPU(const double* Qx, const double* Qy, double* Qxold, double* Qyold, int xsz,
                    int ysz, double dx, af_stream_t stream) {
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(Qx && Qy && Qxold && Qyold, AF_E_INVALID, "UpdateQs: null pointer");
    hipStream_t s = to_stream(stream);
    long long n = (long long)(xsz + 1) * (ysz + 1);
    double dxinv = 1.0 / dx;
    hipLaunchKernelGGL(k_update_qs, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, Qx, Qy,
                       Qxold, Qyold, n, dxinv);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_FloodplainQ_GPU(const double* H, const double* DEM, double* Qx, double* Qy,
                       const double* Qxold, const double* Qyold, int xsz, int ysz, double dx,
                       double dy, double FPn, double DepthThresh, double MaxHflow, double dhlin,
                       double* Tstep, double Qlimfact, double dA, double cfl, double g,
                       double InitTstep, double theta, double nodata, void* workspace,
                       af_stream_t stream) {
    (void)dy; (void)dhlin; (void)Qlimfact; (void)dA;
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(H && DEM && Qx && Qy && Qxold && Qyold && Tstep && workspace, AF_E_INVALID,
               "FloodplainQ: null pointer");
    return floodplain_q(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, FPn, DepthThresh, MaxHflow,
                        Tstep, cfl, g, InitTstep, theta, nodata,
                        reinterpret_cast<unsigned long long*>(workspace), to_stream(stream));
}

int af_FloodplainQ_GPU_noalloc(const double* H, const double* DEM, double* Qx, double* Qy,
                               const double* Qxold, const double* Qyold, int xsz, int ysz,
                               double dx, double dy, double FPn, double DepthThresh,
                               double MaxHflow, double dhlin, double* Tstep, double Qlimfact,
                               double dA, double cfl, double g, double InitTstep, double theta,
                               double nodata, af_stream_t stream) {
    (void)dy; (void)dhlin; (void)Qlimfact; (void)dA;
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(H && DEM && Qx && Qy && Qxold && Qyold && Tstep, AF_E_INVALID,
               "FloodplainQ: null pointer");
    hipStream_t s = to_stream(stream);
    unsigned long long* key = nullptr;
    if ((rc = scratch_ptr(s, &key))) return rc;
    return floodplain_q(H, DEM, Qx, Qy, Qxold, Qyold, xsz, ysz, dx, FPn, DepthThresh, MaxHflow,
                        Tstep, cfl, g, InitTstep, theta, nodata, key, s);
}

int af_BCs_GPU(double* Qx, double* Qy, const double* H, const double* DEM, int xsz, int ysz,
               double dx, double g, double dt, double nodata, int west_free, int east_free,
               int north_free, int south_free, af_stream_t stream) {
    (void)dt;
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(Qx && Qy && H && DEM, AF_E_INVALID, "BCs: null pointer");
    hipStream_t s = to_stream(stream);
    long long total = 2LL * (ysz + 1) + 2LL * (xsz + 1);
    hipLaunchKernelGGL(k_bcs, dim3(blocks_for(total, kThreads)), dim3(kThreads), 0, s, Qx, Qy, H,
                       DEM, xsz, ysz, dx, g, nodata, west_free ? 1 : 0, east_free ? 1 : 0,
                       north_free ? 1 : 0, south_free ? 1 : 0);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_ApplyPointBCs_GPU(double* H, const int32_t* bc_idx, const int32_t* bc_type,
                         const double* bc_value, int num_bcs, int ncols, double dx, double dA,
                         double Tstep, af_stream_t stream) {
    (void)ncols; (void)dx; (void)dA; (void)Tstep;
    AF_REQUIRE(num_bcs >= 0, AF_E_INVALID, "ApplyPointBCs: num_bcs < 0");
    if (num_bcs == 0) return AF_OK;
    AF_REQUIRE(H && bc_idx && bc_type && bc_value, AF_E_INVALID, "ApplyPointBCs: null pointer");
    hipStream_t s = to_stream(stream);
    for (int off = 0; off < num_bcs; off += kSortCap) {
        const int m = std::min(kSortCap, num_bcs - off);
        hipLaunchKernelGGL(k_apply_point_bcs_sorted, dim3(1), dim3(kSortThreads), 0, s, H,
                           bc_idx + off, bc_type + off, bc_value + off, m);
        AF_HIP(hipGetLastError());
    }
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_UpdateH_GPU(double* H, const double* Qx, const double* Qy, int xsz, int ysz, double dA,
                   double Tstep, af_stream_t stream) {
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(H && Qx && Qy, AF_E_INVALID, "UpdateH: null pointer");
    hipStream_t s = to_stream(stream);
    long long n = (long long)xsz * ysz;
    hipLaunchKernelGGL(k_update_h, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, H, Qx, Qy,
                       xsz, ysz, dA, Tstep);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_Rainfall_GPU(double* H, const double* DEM, int xsz, int ysz, double rain_rate,
                    double Tstep, double nodata, af_stream_t stream) {
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    if (rain_rate <= 0.0) return AF_OK;  // kernels.cpp:356
    AF_REQUIRE(H && DEM, AF_E_INVALID, "Rainfall: null pointer");
    hipStream_t s = to_stream(stream);
    long long n = (long long)xsz * ysz;
    double depth = rain_rate * Tstep;
    hipLaunchKernelGGL(k_rainfall, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, H, DEM, n,
                       depth, nodata);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_AddFlowBCs_GPU(double* H, const int32_t* ps_idx, const double* ps_flow, int num_flow_bcs,
                      af_stream_t stream) {
    AF_REQUIRE(num_flow_bcs >= 0, AF_E_INVALID, "AddFlowBCs: count < 0");
    if (num_flow_bcs == 0) return AF_OK;
    AF_REQUIRE(H && ps_idx && ps_flow, AF_E_INVALID, "AddFlowBCs: null pointer");
    hipStream_t s = to_stream(stream);
    for (int off = 0; off < num_flow_bcs; off += kSortCap) {
        const int m = std::min(kSortCap, num_flow_bcs - off);
        hipLaunchKernelGGL(k_add_flow_sorted, dim3(1), dim3(kSortThreads), 0, s, H, ps_idx + off,
                           ps_flow + off, m);
        AF_HIP(hipGetLastError());
    }
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

int af_ComputeVelocity_GPU(double* Vx_out, double* Vy_out, const double* H, const double* DEM,
                           const double* Qx, const double* Qy, int xsz, int ysz, double dx,
                           double nodata, af_stream_t stream) {
    int rc = check_grid(xsz, ysz);
    if (rc) return rc;
    AF_REQUIRE(Vx_out && Vy_out && H && DEM && Qx && Qy, AF_E_INVALID, "Velocity: null pointer");
    hipStream_t s = to_stream(stream);
    long long n = (long long)xsz * ysz;
    hipLaunchKernelGGL(k_velocity, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, Vx_out,
                       Vy_out, H, DEM, Qx, Qy, xsz, ysz, dx, nodata);
    AF_HIP(hipGetLastError());
    AF_HIP(hipStreamSynchronize(s));
    return AF_OK;
}

}  // extern "C"
