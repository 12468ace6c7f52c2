// copy_probe.hip -- calibration program for the HBM traffic counters.
//
// Copies a device buffer of doubles into another with one plain streaming
// kernel (mpd_copy_f64) and prints the size and bandwidth.  Each launch
// reads and writes exactly `bytes`, so the FETCH_SIZE and WRITE_SIZE values
// a counter pass reports for mpd_copy_f64 give the factor from counter units
// to bytes that the stencil's counter readings are scaled by.
//
//   mpd_copy_probe [MiB]   (default 1024; 3 launches)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) mpd_copy_f64(const double2* __restrict__ src, double2* __restrict__ dst,
                                                     long long n2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
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
    const unsigned blocks = 2048;
    for (int it = 0; it < 3; ++it) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL(mpd_copy_f64, dim3(blocks), dim3(256), 0, 0, a, b, n2);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("mpd_copy_f64: %lld bytes read + %lld written (%lld KiB each), %.4f ms, %.1f GB/s\n", bytes,
                    bytes, bytes >> 10, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
