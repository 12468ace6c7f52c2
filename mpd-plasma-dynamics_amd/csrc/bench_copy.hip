// bench_copy.hip -- the copy kernel behind mpd_bench_copy: the measured copy
// ceiling of the GPU's HBM, and the calibration of the HBM traffic counters.
//
// mpd_copy_f64 copies a device buffer of doubles into another, 16 bytes per
// lane per iteration.  The calibration launches use one fixed shape (2048
// workgroups of 256, default stores): each reads and writes exactly the
// buffer size, so the FETCH_SIZE and WRITE_SIZE a counter pass reports for
// them give the factor from counter units to bytes.  The sweep over
// workgroup counts and store policy (default / non-temporal) gives the best
// copy bandwidth.  No counterpart in the reference (a benchmark helper).
#include "solver_core.h"

namespace {

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

void Launch(bool nt, unsigned blocks, const double2* a, double2* b, long long n2, hipStream_t s) {
    if (nt) hipLaunchKernelGGL(mpd_copy_f64<true>, dim3(blocks), dim3(256), 0, s, a, b, n2);
    else hipLaunchKernelGGL(mpd_copy_f64<false>, dim3(blocks), dim3(256), 0, s, a, b, n2);
}

struct Buffers {
    double2* a = nullptr;
    double2* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = nullptr;
    ~Buffers() {
        if (s) (void)hipStreamSynchronize(s);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

float Elapsed(Buffers& m) {
    HIPCHK(hipEventSynchronize(m.e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m.e0, m.e1));
    return ms;
}

}  // namespace

extern "C" int mpd_bench_copy(long long mib, int calib_launches, double* calib_ms, double* ceiling_gbs,
                              int* best_blocks, int* best_nt) {
    return Guard([&] {
        if (mib <= 0 || mib > 65536) throw ArgError("size must be 1..65536 MiB");
        if (calib_launches < 0 || calib_launches > 64) throw ArgError("calib_launches must be 0..64");
        const long long bytes = mib << 20;
        const long long n2 = bytes / 16;
        Buffers m;
        HIPCHK(hipStreamCreate(&m.s));
        HIPCHK(hipMalloc(&m.a, bytes));
        HIPCHK(hipMalloc(&m.b, bytes));
        HIPCHK(hipMemsetAsync(m.a, 0, bytes, m.s));
        HIPCHK(hipMemsetAsync(m.b, 0, bytes, m.s));
        HIPCHK(hipEventCreate(&m.e0));
        HIPCHK(hipEventCreate(&m.e1));
        double cal = 0.0;
        for (int it = 0; it < calib_launches; ++it) {
            HIPCHK(hipEventRecord(m.e0, m.s));
            Launch(false, 2048, m.a, m.b, n2, m.s);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(m.e1, m.s));
            cal += Elapsed(m);
        }
        if (calib_ms) *calib_ms = calib_launches ? cal / calib_launches : 0.0;
        double best = 0.0;
        int bb = 0, bnt = 0;
        const unsigned shapes[] = {1024, 2048, 4096, 8192, 16384};
        for (int nt = 0; nt < 2; ++nt) {
            for (unsigned blocks : shapes) {
                const int reps = 10;
                for (int w = 0; w < 2; ++w) Launch(nt, blocks, m.a, m.b, n2, m.s);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(m.e0, m.s));
                for (int r = 0; r < reps; ++r) Launch(nt, blocks, m.a, m.b, n2, m.s);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(m.e1, m.s));
                const double gbs = 2.0 * (double)bytes * reps / (Elapsed(m) * 1e-3) / 1e9;
                if (gbs > best) {
                    best = gbs;
                    bb = (int)blocks;
                    bnt = nt;
                }
            }
        }
        if (ceiling_gbs) *ceiling_gbs = best;
        if (best_blocks) *best_blocks = bb;
        if (best_nt) *best_nt = bnt;
    });
}
