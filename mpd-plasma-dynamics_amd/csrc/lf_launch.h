// lf_launch.h -- host-side launchers of the kernels in lf_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "lf_device.h"

namespace mpd {

enum StencilKind { STENCIL_SIMPLE = 0, STENCIL_TILED = 1 };

// Row layout of every pitched field buffer.  The columns m = 0 .. M_max + 2
// of a row are cut into strips of kStripCols consecutive columns; strip s
// occupies the kSlotCols = 64 doubles (512 bytes, four 128-byte lines) from
// s * 64, its last two doubles unused.  The tiled stencil gives strip s to one
// wavefront whose 64 lanes hold columns 62 s - 1 .. 62 s + 62: the two outer
// lanes read the edge columns of the neighbouring strips in the same row load
// in which the neighbouring wavefronts read them, and every store covers one
// whole slot (the two outer lanes write the unused doubles).
constexpr int kStripCols = 62;
constexpr int kSlotCols = 64;

__host__ __device__ inline int phys_col(int m) { return (m / kStripCols) * kSlotCols + m % kStripCols; }
inline int layout_pitch(int cols) { return (cols + kStripCols - 1) / kStripCols * kSlotCols; }

struct StencilCfg {
    int kind = STENCIL_TILED;
    bool fast_div = true;
    int tile_rows = 0;  // rows per wave of the tiled kernel (0: chosen from the grid)
    int order = 0;      // tile order: 0 strip-fastest, 1 row bands per XCD, 2 stacked tiles, 3 strip groups per XCD (see lf_stencil_tiled)
    // Raw buffer addressing when every state is below 2^31 bytes (false:
    // always the plain-pointer path, which larger states take).
    bool buffers = true;
};

// Central update (+ wave-speed epilogue when speed).  loop: read dt and the
// done flag from the device loop state instead of dt_explicit.
hipError_t launch_stencil(const StepArgs& a, const StencilCfg& c, double dt_explicit, int loop, int speed,
                          hipStream_t s);
// Boundary faces; with loop, a one-thread kernel after them finishes the
// step's loop state (rescan test, dt bookkeeping).
hipError_t launch_bc(const StepArgs& a, double dt_explicit, int loop, int speed, hipStream_t s);
// Exact max-speed rescan + dt bookkeeping when the boundary kernel flagged
// that the step's candidate max is not provably the max (loop only).
hipError_t launch_speed_fallback(const StepArgs& a, hipStream_t s);
// SolutionChange (+ SavePrev when save).  loop: run only on a check step.
hipError_t launch_residual(const StepArgs& a, double* prev, double* partial, int loop, int check_frequency,
                           double threshold, int save, hipStream_t s);
// Workgroups of one instance's SolutionChange reduction: every launch of it
// (single solver, batch, rescan path) partitions the cells the same way, so
// an instance's change has the same bits however it is run.
__host__ __device__ inline int residual_blocks(int L, int Mn) {
    const long long b = ((long long)L * Mn + 255) / 256;
    return (int)(b < 1 ? 1 : (b < 1024 ? b : 1024));
}

// Batched launches over many instances (work lists built by the host; args
// holds two StepArgs per instance, icur the start buffer of each instance).
// grid: workgroups of the stencil launch (<= nwork; fewer makes the grid
// persistent over the list).
hipError_t launch_stencil_batch(const StepArgs* args, const int* icur, const BatchBlock* work, int nwork, int grid,
                                int flip, bool fast_div, bool buf, double dt_explicit, int loop, hipStream_t s);
hipError_t launch_bc_batch(const StepArgs* args, const int* icur, const BatchBlock* work, int nwork, int flip,
                           int img_words, int area_slots, hipStream_t s);
// The step's end (rescan test, dt bookkeeping) of every instance that took it.
hipError_t launch_finish_batch(const StepArgs* args, const int* icur, int n, int flip, hipStream_t s);
hipError_t launch_residual_batch(const StepArgs* args, const int* icur, const BatchAux* aux, const BatchBlock* work,
                                 int nwork, int flip, hipStream_t s);
// The persistent batched loop of small instances: one 256-thread workgroup
// per work entry {instance, 0, its tile count, tile rows, strips}, up to K
// steps of that instance per launch.  LDS: the largest program image
// (img_words) and four work areas of area_slots slots per lane.
size_t persist_lds_bytes(int img_words, int area_slots);
hipError_t launch_batch_persist(const StepArgs* args, const int* icur, const BatchAux* aux, const BatchBlock* work,
                                int nwork, int K, bool fast_div, bool buf, int img_words, int area_slots,
                                hipStream_t s);
// Tiles of one instance for tile height H (strips of central columns x row chunks).
inline int stencil_strips(int Mn) { return (Mn - 1) / kStripCols + 1; }
inline int stencil_tiles(int L, int Mn, int H) { return stencil_strips(Mn) * ((L + H - 1) / H); }
inline int default_tile_rows() { return 8; }
hipError_t launch_max_speed(const StepArgs& a, unsigned long long* out, hipStream_t s);
hipError_t launch_phys_fields(const StepArgs& a, double* out, hipStream_t s);
// Physical fields of rows r0 .. r0 + nrows - 1 (10 fields of nrows pitched
// rows each, field k at out + k * nrows * pitch).
hipError_t launch_phys_rows(const StepArgs& a, int r0, int nrows, double* out, hipStream_t s);
hipError_t launch_integrals(const StepArgs& a, double* rowsum, double* out, hipStream_t s);
hipError_t launch_outlet(const StepArgs& a, int M_max, double* out, hipStream_t s);
hipError_t launch_radial_sym(const StepArgs& a, unsigned long long* out, hipStream_t s);
hipError_t launch_fill(double* p, long long n, double v, hipStream_t s);
// Physical values of cell (l, m) into out[0..9].
hipError_t launch_probe(const StepArgs& a, int l, int m, double* out, hipStream_t s);
// y[i] = div_recip(2 d[i]) for i < n.
hipError_t launch_recip_table(const double* d, double* y, int n, hipStream_t s);

}  // namespace mpd
