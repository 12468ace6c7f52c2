// noalloc.hip -- mpd_download_noalloc: the state download of mpd_download
// without any device allocation, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// mpd_download computes the 10 physical arrays of the whole state into a
// staging buffer it allocates for the call (10 x rows x pitch doubles).
// This variant stages them through `scratch`, a file-scope device array of
// kScratchDoubles doubles (16 MiB; every GPU the library runs on holds its
// own copy until the process ends), a chunk of whole rows at a time, each
// chunk copied out before the next is computed.  The state's conserved
// arrays are copied straight from the state buffer, as mpd_download does.
//
// In host code the name `scratch` only stands for the array: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device
// (with that device current) at the first call on it, and only that address
// is passed to kernels and copies.  Calls for solvers on one GPU share that
// GPU's array; the library serialises them with a per-device lock, so
// concurrent calls from threads of one process do not race, but another
// process's library copy has its own array.
#include <mutex>

#include "solver_core.h"

namespace {

constexpr long long kScratchDoubles = 2LL * 1024 * 1024;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::once_flag once;
    hipError_t err = hipSuccess;
    double* ptr = nullptr;
    std::mutex lock;  // one call at a time uses this device's array
};
DeviceScratch g_scratch[kMaxDevices];

// Makes `dev` current for the scope, restoring the previous device after.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        HIPCHK(hipGetDevice(&prev));
        if (prev != dev) HIPCHK(hipSetDevice(dev));
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// The address of `scratch` on device dev (current), looked up once.
double* ScratchAddress(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw ArgError("device index out of range for the no-allocation scratch");
    DeviceScratch& d = g_scratch[dev];
    std::call_once(d.once, [&] {
        void* p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        d.ptr = static_cast<double*>(p);
    });
    if (d.err != hipSuccess || !d.ptr)
        throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(d.err));
    return d.ptr;
}

}  // namespace

extern "C" int mpd_download_noalloc(mpd_solver* s, double* const* u0, double* const* phys) {
    return Guard([&] {
        Need(s);
        const long long row_doubles = 10LL * s->pitch;
        if (phys && row_doubles > kScratchDoubles)
            throw ArgError("mpd_download_noalloc: one grid row of the 10 physical fields needs " +
                           std::to_string(row_doubles) + " doubles, more than the " + std::to_string(kScratchDoubles) +
                           "-double no-allocation scratch (M_max too large; use mpd_download)");
        DeviceScope scope(s->device);
        double* scr = ScratchAddress(s->device);
        std::lock_guard<std::mutex> hold(g_scratch[s->device].lock);
        HIPCHK(hipStreamSynchronize(s->stream));
        if (u0) s->DownloadFields(s->state[s->cur], u0, 8);
        if (!phys) return;
        if (s->step1) {
            // Before the first step the physical arrays are the uploaded ones.
            s->DownloadFields(s->phys_ic, phys, 10);
            return;
        }
        const int chunk = (int)std::min<long long>(s->rows, kScratchDoubles / row_doubles);
        std::vector<double> stage((size_t)chunk * s->pitch);
        const mpd::StepArgs a = s->Args(s->cur);
        for (int r0 = 0; r0 < s->rows; r0 += chunk) {
            const int n = std::min(chunk, s->rows - r0);
            HIPCHK(mpd::launch_phys_rows(a, r0, n, scr, s->stream));
            for (int k = 0; k < 10; ++k) {
                if (!phys[k]) continue;
                HIPCHK(hipMemcpyAsync(stage.data(), scr + (long long)k * n * s->pitch,
                                      (size_t)n * s->pitch * sizeof(double), hipMemcpyDeviceToHost, s->stream));
                HIPCHK(hipStreamSynchronize(s->stream));
                for (int i = 0; i < n; ++i) {
                    const double* in = &stage[(size_t)i * s->pitch];
                    double* row = phys[k] + (size_t)(r0 + i) * s->cols;
                    mpd_solver::ForRuns(0, s->cols - 1, [&](int m, int pc, int cnt) {
                        std::memcpy(row + m, in + pc, sizeof(double) * cnt);
                    });
                }
            }
        }
    });
}
