// hbm_probe.hip -- the HBM bandwidth the fused step's access pattern can reach.
//
// The fused step reads four arrays (z, H, Qx, Qy) and writes three (H, Qx, Qy)
// in row-major rows of pitch P, each wavefront marching a 256-column strip
// down R rows with one 16-byte access per lane and array (the state layout of
// fused.hip at 16384 columns).  These kernels make the same accesses with no
// stencil and no flux arithmetic, so their rate is the ceiling of the step's
// memory pattern on the machine at hand:
//   step       4 reads + 3 non-temporal stores (the step's form)
//   halo       the same plus the step's four strip-edge column loads per row
//   nt-loads   the step's form with non-temporal loads as well
//   read       the four reads only
//   copy       one read and one non-temporal store
// Usage: af_hbm_probe [n] [rows] [reps].  Prints one line per pattern.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
            exit(1);                                                                   \
        }                                                                              \
    } while (0)

typedef float f4 __attribute__((ext_vector_type(4)));

struct Arrays {
    const float *z, *h, *qx, *qy;
    float *ho, *qxo, *qyo;
};

enum { kStep = 0, kHalo = 1, kNtLoads = 2, kRead = 3, kCopy = 4 };

// blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of strips, as the step kernel does
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x, q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_pattern(Arrays a, int n, long long P, int R) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nstrips = n / 256;
    const int tile = xcd_block() * 4 + wv;
    if (tile >= nstrips * ((n + R - 1) / R)) return;
    const int strip = tile % nstrips, seg = tile / nstrips;
    const int j0 = seg * R, j1 = min(j0 + R, n);
    const int i0 = strip * 256, c0 = i0 + lane * 4;
    int hc = (lane == 0) ? i0 - 1 : i0 + 256;
    if (hc < 0 || hc >= n) hc = i0;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    float hacc = 0.f;
    for (int j = j0; j < j1; j++) {
        const long long o = (long long)j * P + c0;
        if (MODE == kCopy) {
            const f4 h = *(const f4*)(a.h + o);
            __builtin_nontemporal_store(h * 1.0001f, (f4*)(a.ho + o));
            continue;
        }
        f4 z, h, qx, qy;
        if (MODE == kNtLoads) {
            z = __builtin_nontemporal_load((const f4*)(a.z + o));
            h = __builtin_nontemporal_load((const f4*)(a.h + o));
            qx = __builtin_nontemporal_load((const f4*)(a.qx + o));
            qy = __builtin_nontemporal_load((const f4*)(a.qy + o));
        } else {
            z = *(const f4*)(a.z + o);
            h = *(const f4*)(a.h + o);
            qx = *(const f4*)(a.qx + o);
            qy = *(const f4*)(a.qy + o);
        }
        if (MODE == kHalo) {
            const long long oh = (long long)j * P + hc;
            hacc += a.z[oh] + a.h[oh] + a.qx[oh] + a.qy[oh];
        }
        if (MODE == kRead) {
            acc += z + h + qx + qy;
            continue;
        }
        const f4 hn = h + z * 1e-7f + qx * 1e-7f + qy * 1e-7f + hacc * 1e-9f;
        __builtin_nontemporal_store(hn, (f4*)(a.ho + o));
        __builtin_nontemporal_store(qx * 0.5f + z * 1e-7f, (f4*)(a.qxo + o));
        __builtin_nontemporal_store(qy * 0.5f + h * 1e-7f, (f4*)(a.qyo + o));
        acc += hn;
    }
    if (acc.x == 12345.0f) a.ho[0] = acc.y;  // keeps the read-only sums live
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 16384;
    const int R = argc > 2 ? atoi(argv[2]) : 64;
    const int reps = argc > 3 ? atoi(argv[3]) : 2;
    if (n < 256 || n % 256 != 0 || R < 1) {
        fprintf(stderr, "n must be a positive multiple of 256, rows >= 1\n");
        return 2;
    }
    // the step's pitch for n columns: an odd multiple of 1 KiB
    long long row = ((long long)(n + 1) * 4 + 1023) / 1024 * 1024;
    if (((row >> 10) & 1) == 0) row += 1024;
    const long long P = row / 4;
    const size_t bytes = (size_t)n * P * 4;
    float* buf[7];
    for (int i = 0; i < 7; i++) {
        CK(hipMalloc(&buf[i], bytes));
        CK(hipMemset(buf[i], 0, bytes));
    }
    Arrays a{buf[0], buf[1], buf[2], buf[3], buf[4], buf[5], buf[6]};
    const int ntiles = (n / 256) * ((n + R - 1) / R);
    const int blocks = (ntiles + 3) / 4;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    auto run = [&](auto kern, double bytes_per_cell, const char* name) {
        for (int w = 0; w < 3; w++)
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, a, n, P, R);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        const int it = 20;
        CK(hipEventRecord(e0));
        for (int w = 0; w < it; w++)
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, a, n, P, R);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= it;
        printf("%-9s %8.4f ms %8.1f GB/s  (%.0f B/cell, n=%d, rows=%d, pitch=%lld)\n", name, ms,
               bytes_per_cell * (double)n * n / ms / 1e6, bytes_per_cell, n, R, P);
        fflush(stdout);
    };
    for (int r = 0; r < reps; r++) {
        run(k_pattern<kStep>, 28, "step");
        run(k_pattern<kHalo>, 28, "halo");
        run(k_pattern<kNtLoads>, 28, "nt-loads");
        run(k_pattern<kRead>, 16, "read");
        run(k_pattern<kCopy>, 8, "copy");
    }
    for (int i = 0; i < 7; i++) CK(hipFree(buf[i]));
    return 0;
}
