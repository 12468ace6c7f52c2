// copy_probe.hip -- calibration program for the HBM traffic counters and the
// measured copy ceiling.
//
// mpd_copy_f64 copies a device buffer of doubles into another, 16 bytes per
// lane per iteration.  The first three launches use one fixed shape (2048
// workgroups of 256, default stores): each reads and writes exactly `bytes`,
// so the FETCH_SIZE and WRITE_SIZE a counter pass reports for them give the
// factor from counter units to bytes.  Then a sweep over workgroup counts
// and store policy (default / non-temporal) prints the best copy bandwidth,
// the practical HBM ceiling of this GPU.
//
//   mpd_copy_probe [MiB]   (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

template <bool kNT>
__global__ void __launch_bounds__(256) mpd_copy_f64(const double2* __restrict__ src, double2* __restrict__ dst,
                                                     long long n2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const double2 v = src[i];
        if (kNT) {
            __builtin_nontemporal_store(v.x, &dst[i].x);
            __builtin_nontemporal_store(v.y, &dst[i].y);
        } else {
            dst[i] = v;
        }
    }
}

#define CHECK(x)                                                                      \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

int main(int argc, char** argv) {
    const long long mib = argc > 1 ? std::atoll(argv[1]) : 1024;
    if (mib <= 0 || mib > 65536) {
        std::fprintf(stderr, "size must be 1..65536 MiB\n");
        return 2;
    }
    const long long bytes = mib << 20;
    const long long n2 = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int it = 0; it < 3; ++it) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL(mpd_copy_f64<false>, dim3(2048), dim3(256), 0, 0, a, b, n2);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("mpd_copy_f64: %lld bytes read + %lld written (%lld KiB each), %.4f ms, %.1f GB/s\n", bytes,
                    bytes, bytes >> 10, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    double best = 0.0;
    const char* best_policy = "";
    unsigned best_blocks = 0;
    const unsigned shapes[] = {1024, 2048, 4096, 8192, 16384};
    for (int nt = 0; nt < 2; ++nt) {
        for (unsigned blocks : shapes) {
            const int reps = 10;
            for (int w = 0; w < 2; ++w) {
                if (nt) hipLaunchKernelGGL(mpd_copy_f64<true>, dim3(blocks), dim3(256), 0, 0, a, b, n2);
                else hipLaunchKernelGGL(mpd_copy_f64<false>, dim3(blocks), dim3(256), 0, 0, a, b, n2);
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e0));
            for (int r = 0; r < reps; ++r) {
                if (nt) hipLaunchKernelGGL(mpd_copy_f64<true>, dim3(blocks), dim3(256), 0, 0, a, b, n2);
                else hipLaunchKernelGGL(mpd_copy_f64<false>, dim3(blocks), dim3(256), 0, 0, a, b, n2);
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double gbs = 2.0 * bytes * reps / (ms * 1e-3) / 1e9;
            if (gbs > best) {
                best = gbs;
                best_policy = nt ? "non-temporal" : "default";
                best_blocks = blocks;
            }
        }
    }
    std::printf("copy ceiling: %.1f GB/s (%u workgroups of 256, %s stores, %lld MiB per buffer)\n", best,
                best_blocks, best_policy, mib);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
