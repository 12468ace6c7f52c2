// lf_kernels.hip -- gfx950 kernels of the explicit Lax-Friedrichs step.
//
//   lf_stencil_*      the fused flux-plus-update stencil over the 8 conserved
//                     variables with the axisymmetric sources
//                     (Solver::ComputeCentralUpdate, src/solver.cpp:59-153),
//                     physical fields recomputed in registers from u0, and the
//                     max wave speed of the new state as an epilogue
//                     (Diagnostics::MaxWaveSpeed, src/diagnostics.cpp:14-43).
//   lf_bc_ring        the four faces L_LO, M_HI, M_LO, L_HI in the reference
//                     order (src/solver.cpp:29-32, src/face_bc.cpp,
//                     src/bcs/per_field_bc.cpp), one thread per boundary cell,
//                     plus the dt update of the time loop (src/main.cpp:221-231).
//   lf_residual       SolutionChange + SavePrev (src/diagnostics.cpp:71-106,
//                     src/fields.cpp:59-74).
//   reductions        max wave speed, conserved integrals, outlet integrals,
//                     radial-symmetry metric, physical-field download.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lf_device.h"
#include "lf_launch.h"

namespace mpd {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ const double* fld(const double* base, int64_t fstride, int k) {
    return base + (int64_t)k * fstride;
}

__device__ __forceinline__ int64_t at(const StepArgs& a, int l, int m) {
    return (int64_t)l * a.pitch + phys_col(m);
}

// Start-of-step physical value set at an interior cell: f(u0), or the IC
// arrays on the first step (src/main.cpp:144-152 sets them before step 1).
__device__ __forceinline__ void start_phys(const StepArgs& a, int l, int m, double* ph) {
    if (a.step1) {
        for (int k = 0; k < 10; ++k) ph[k] = fld(a.phys_in, a.fstride, k)[at(a, l, m)];
    } else {
        double u[8];
        for (int k = 0; k < 8; ++k) u[k] = fld(a.u_in, a.fstride, k)[at(a, l, m)];
        phys10_from_u(u, cell_r(a, l, m), a.gm1, ph);
    }
}

// start_phys and s0_phys (below) split in two, so that the boundary kernel
// can issue the loads of several cells before it waits for any: where the
// values of cell (l, m) live (8 conserved values at p + k stride, or on the
// first step 10 physical values), the loads, and the conversion.
struct CellSrc {
    const double* p;
    int64_t stride;
    bool phys;
};

__device__ __forceinline__ CellSrc start_src(const StepArgs& a, int l, int m) {
    if (a.step1) return {a.phys_in + at(a, l, m), a.fstride, true};
    return {a.u_in + at(a, l, m), a.fstride, false};
}

// The physical state every boundary-condition read starts from (the state
// after UpdateCentralPhysical, src/solver.cpp:155-159): central cells hold
// f(u_new), every other interior cell the start-of-step physical values.
__device__ __forceinline__ void s0_phys(const StepArgs& a, int l, int m, double* ph) {
    if (m >= 2 && m <= a.Mn - 1) {
        double u[8];
        if (l == 1 || l == a.L) {
            const double* rb = a.rowbuf + (int64_t)(l == 1 ? 0 : 8) * a.pitch;
            for (int k = 0; k < 8; ++k) u[k] = rb[(int64_t)k * a.pitch + phys_col(m)];
        } else {
            for (int k = 0; k < 8; ++k) u[k] = fld(a.u_out, a.fstride, k)[at(a, l, m)];
        }
        phys10_from_u(u, cell_r(a, l, m), a.gm1, ph);
    } else {
        start_phys(a, l, m, ph);
    }
}

__device__ __forceinline__ CellSrc s0_src(const StepArgs& a, int l, int m) {
    if (m >= 2 && m <= a.Mn - 1) {
        if (l == 1 || l == a.L) return {a.rowbuf + (int64_t)(l == 1 ? 0 : 8) * a.pitch + phys_col(m), a.pitch, false};
        return {a.u_out + at(a, l, m), a.fstride, false};
    }
    return start_src(a, l, m);
}

__device__ __forceinline__ void src_load(const CellSrc& s, double* v) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = s.p[k * s.stride];
    if (s.phys) {
        v[8] = s.p[8 * s.stride];
        v[9] = s.p[9 * s.stride];
    }
}

__device__ __forceinline__ void src_phys(const CellSrc& s, const double* v, double r, double gm1, double* ph) {
    if (s.phys) {
#pragma unroll
        for (int k = 0; k < 10; ++k) ph[k] = v[k];
    } else {
        phys10_from_u(v, r, gm1, ph);
    }
}

// ------------------------------------------------------------- the stencil

// Loads the stencil record of cell (l, m) from u0 (or the IC arrays on the
// first step).  Ghost rows l = 0 and l = L+1 hold u0 = 0 and physical = 0 in
// the reference (never written, zero-initialised, src/array2d.cpp:30).
template <bool kStep1>
__device__ __forceinline__ void load_rec(const StepArgs& a, int l, int m, Rec& c) {
    const int64_t i = at(a, l, m);
    for (int k = 0; k < 8; ++k) c.u[k] = fld(a.u_in, a.fstride, k)[i];
    c.r = cell_r(a, l, m);
    if (kStep1) {
        const double* ph = a.phys_in;
        c.rho = fld(ph, a.fstride, PH_RHO)[i];
        c.vz = fld(ph, a.fstride, PH_VZ)[i];
        c.vr = fld(ph, a.fstride, PH_VR)[i];
        c.vphi = fld(ph, a.fstride, PH_VPHI)[i];
        c.p = fld(ph, a.fstride, PH_P)[i];
        c.P = fld(ph, a.fstride, PH_PT)[i];
        c.Hz = fld(ph, a.fstride, PH_HZ)[i];
        c.Hr = fld(ph, a.fstride, PH_HR)[i];
        c.Hphi = fld(ph, a.fstride, PH_HPHI)[i];
    } else if (l == 0 || l == a.L + 1) {
        c.rho = c.vz = c.vr = c.vphi = c.p = c.P = c.Hz = c.Hr = c.Hphi = 0.0;
    } else {
        rec_phys(c, a.gm1);
    }
}

// A loop step does nothing once the loop has ended (done) or while the last
// step waits for its exact wave-speed rescan (fallback): the host runs the
// rescan, which finishes that step, and launches the remaining steps again.
__device__ __forceinline__ bool loop_idle(const StepArgs& a) {
    return (a.st->done | a.st->fallback | a.st->range_stop) != 0;
}

__device__ __forceinline__ double step_dt(const StepArgs& a, double dt_explicit, int loop) {
    return loop ? a.st->dt_use : dt_explicit;
}

// One thread per central cell, every neighbour record loaded and recomputed
// independently (5x redundant physical recompute): the plain reference-shaped
// kernel, used for the first step and as a cross-check of the tiled kernel.
template <bool kStep1, bool kSpeed, class Div>
__global__ void __launch_bounds__(256) lf_stencil_simple(StepArgs a, double dt_explicit, int loop) {
    if (loop && loop_idle(a)) return;
    const int m = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int l = 1 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const double dt = step_dt(a, dt_explicit, loop);
    uint64_t key = 0;
    bool rbad = false;
    if (m <= a.Mn - 1 && l <= a.L) {
        Rec N, S, E, W, C;
        load_rec<kStep1>(a, l + 1, m, N);
        load_rec<kStep1>(a, l - 1, m, S);
        load_rec<kStep1>(a, l, m + 1, E);
        load_rec<kStep1>(a, l, m - 1, W);
        load_rec<kStep1>(a, l, m, C);
        const double dr_l = a.dr[l];
        const double dr2 = 2 * dr_l;
        double o[8];
        lf_cell<Div>(N, S, E, W, C, dt, a.dz2, div_recip(a.dz2), dr2, div_recip(dr2), o);
        if (l == 1 || l == a.L) {
            double* rb = a.rowbuf + (int64_t)(l == 1 ? 0 : 8) * a.pitch;
            for (int k = 0; k < 8; ++k) rb[(int64_t)k * a.pitch + phys_col(m)] = o[k];
        } else {
            const int64_t i = at(a, l, m);
            for (int k = 0; k < 8; ++k) a.u_out[(int64_t)k * a.fstride + i] = o[k];
            if (kSpeed) key = speed_key(wave_speed_u(o, C.r, a.gamma, a.gm1));
        }
        rbad = !split_div_safe(o);
    }
    if (kSpeed) {
        key = wave_max_u64(key);
        if ((threadIdx.x & 63) == 0 && key != 0) atomicMax((unsigned long long*)&a.st->speed_bits, (unsigned long long)key);
    }
    flag_range(a.st, rbad);
}

// A per-row geometry value (row l wave-uniform) read through the constant
// address space: the geometry arrays are not written during a step, and a
// uniform load from that space is a scalar load.  Scalar loads have their
// own completion counter; a vector load waited for in the march would also
// wait for all of the wave's older stores (one counter for vector loads and
// stores).
__device__ __forceinline__ double row_value(const double* p, int l) {
    using CDouble = const __attribute__((address_space(4))) double;
    return ((CDouble*)p)[l];
}

// Record of row l for the marching kernel from its 8 conserved values; zero
// physical values on the ghost rows and beyond the last interior column.
__device__ __forceinline__ void row_rec(const StepArgs& a, int l, int m, const double* raw, Rec& c) {
    for (int k = 0; k < 8; ++k) c.u[k] = raw[k];
    // cell_r (Grid::Build, src/grid.cpp:31-37) with scalar loads of rin, rout.
    const double frac = (double)(m - 1) * a.dy;
    c.r = (1.0 - frac) * row_value(a.rin, l) + frac * row_value(a.rout, l);
    rec_phys_shared(c, a.gm1);
    // Selects, not a branch around rec_phys_shared: two definitions of the
    // record merged at a join would put it in private memory.
    const bool zero = (l == 0 || l == a.L + 1 || m < 1 || m > a.Mn);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(zero) != 0, 0)) {
        c.rho = zero ? 0.0 : c.rho;
        c.vz = zero ? 0.0 : c.vz;
        c.vr = zero ? 0.0 : c.vr;
        c.vphi = zero ? 0.0 : c.vphi;
        c.p = zero ? 0.0 : c.p;
        c.P = zero ? 0.0 : c.P;
        c.Hz = zero ? 0.0 : c.Hz;
        c.Hr = zero ? 0.0 : c.Hr;
        c.Hphi = zero ? 0.0 : c.Hphi;
    }
}

// State-buffer access of the marching kernel.  With kBuf, raw buffer
// instructions: a scalar descriptor for the buffer, the row and field offset
// in a scalar register and this lane's column offset in a 32-bit vector
// register, so no per-lane 64-bit address arithmetic (the whole 8-field
// buffer must stay below 2^31 bytes).  Without kBuf, plain pointers.  Loads
// and stores have their own lane offsets (see lf_stencil_tiled).
// Cache policy of the state stores: non-temporal (the new state is not read
// again in this launch; measured 5% faster than the default policy).
constexpr int kStoreNT = 2;

template <bool kBuf>
struct StateIO {
    __amdgpu_buffer_rsrc_t rin, rout, rrow;
    const double* pin;
    double* pout;
    int64_t row_bytes, field_bytes;
    uint32_t ld_off, st_off;

    __device__ __forceinline__ void init(const StepArgs& a, uint32_t ld, uint32_t st) {
        ld_off = ld;
        st_off = st;
        row_bytes = (int64_t)a.pitch * 8;
        field_bytes = a.fstride * 8;
        pin = a.u_in;
        pout = a.u_out;
        if (kBuf) {
            rin = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(a.u_in), 0, -1, 0x00020000);
            // The store descriptors carry their buffer's exact size (8 fields;
            // 16 rows for the row buffer): a lane that must not store gets
            // the offset kNoStore, past the end, and the hardware drops its
            // store (a raw buffer access at or beyond num_records is not
            // performed).  Every lane then executes the stores, with no
            // branch around them.
            rout = __builtin_amdgcn_make_buffer_rsrc(a.u_out, 0, (int)(8 * a.fstride * 8), 0x00020000);
            rrow = __builtin_amdgcn_make_buffer_rsrc(a.rowbuf, 0, (int)(16 * (int64_t)a.pitch * 8), 0x00020000);
        }
    }
    __device__ __forceinline__ void load_row(int l, double* raw) const {
        if (kBuf) {
            const int so = (int)(l * row_bytes);
            for (int k = 0; k < 8; ++k) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(rin, ld_off, so + (int)(k * field_bytes), 0);
                raw[k] = __builtin_bit_cast(double, v);
            }
        } else {
            const char* row = reinterpret_cast<const char*>(pin) + l * row_bytes + ld_off;
            for (int k = 0; k < 8; ++k) raw[k] = *reinterpret_cast<const double*>(row + k * field_bytes);
        }
    }
    // Offset of a lane that does not store: at or beyond any buffer size the
    // raw-buffer path allows (every buffer below 2^31 bytes).
    static constexpr uint32_t kNoStore = 0x80000000u;
    // Row l of the march's output from every lane, raw buffers only: rows 1
    // and L (rb) go to the row buffer (field k of row 1 at its row k, of
    // row L at row 8 + k), the others to the new state; lanes with `on`
    // false store nothing.  Descriptor and offsets are wave-uniform selects,
    // so every row issues the same eight stores with no branch around them.
    __device__ __forceinline__ void store_out(int l, bool rb, const double* o, bool on) const {
        const __amdgpu_buffer_rsrc_t r = rb ? rrow : rout;
        const int base = rb ? (l == 1 ? 0 : 8) * (int)row_bytes : l * (int)row_bytes;
        const int step = rb ? (int)row_bytes : (int)field_bytes;
        const uint32_t vo = on ? st_off : kNoStore;
        using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(rin, 0, 0, 0));
        for (int k = 0; k < 8; ++k)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, o[k]), r, vo, base + k * step, kStoreNT);
    }
    __device__ __forceinline__ void store_row(int l, const double* o) const {
        if (kBuf) {
            const int so = (int)(l * row_bytes);
            using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(rin, 0, 0, 0));
            for (int k = 0; k < 8; ++k)
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, o[k]), rout, st_off,
                                                      so + (int)(k * field_bytes), kStoreNT);
        } else {
            char* row = reinterpret_cast<char*>(pout) + l * row_bytes + st_off;
            for (int k = 0; k < 8; ++k)
                __builtin_nontemporal_store(o[k], reinterpret_cast<double*>(row + k * field_bytes));
        }
    }
};

// Lane exchange by DPP wave shifts, one v_mov_b32_dpp per 32-bit half:
// DPP_WAVE_SHR1 gives lane i the value of lane i-1, DPP_WAVE_SHL1 the value
// of lane i+1 (the lane without a source gets 0; it is an edge lane whose
// result is not used).
#define DPP_WAVE_SHL1 0x130
#define DPP_WAVE_SHR1 0x138
template <int kCtrl>
__device__ __forceinline__ double dpp_shift(double v) {
    const long long vi = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)vi, kCtrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(vi >> 32), kCtrl, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Neighbour access of lf_cell_rad: each lane computes the radial record
// entry of its own cell and the DPP shift brings the E / W neighbour's.
struct LaneNb {
    const Rec& c;
    __device__ __forceinline__ double E(int i) const { return dpp_shift<DPP_WAVE_SHL1>(radial_entry(c, i)); }
    __device__ __forceinline__ double W(int i) const { return dpp_shift<DPP_WAVE_SHR1>(radial_entry(c, i)); }
};

// Single-precision estimate of the wave speed v + c_s + c_a of conserved
// values u at radius r (gg1 = gamma (gamma - 1), so c_s^2 = gg1 e).  Its
// relative error is a few 2^-23 (hardware reciprocal and square root, one
// ulp each); SPEED_EST_TOL bounds it with a wide margin.  NaN or infinity
// whenever an input is non-finite or out of single range.
__device__ __forceinline__ float speed_estimate(const double* u, double r, float gg1) {
    const float fu0 = (float)u[0], fr = (float)r;
    const float iu = __builtin_amdgcn_rcpf(fu0);
    const float ir = __builtin_amdgcn_rcpf(fr);
    const float vz = (float)u[1] * iu, vr = (float)u[2] * iu, e = (float)u[4] * iu;
    const float hz = (float)u[6] * ir, hr = (float)u[7] * ir, hp = (float)u[5];
    const float cs = __builtin_amdgcn_sqrtf(gg1 * e);
    const float ca = __builtin_amdgcn_sqrtf((hz * hz + hr * hr + hp * hp) * (fr * iu));
    const float v = __builtin_amdgcn_sqrtf(vz * vz + vr * vr);
    return v + cs + ca;
}

// Threshold of the exact wave-speed evaluation in the loop: only cells whose
// estimate reaches SPEED_CAND_FRAC of T = (1 - SPEED_DROP) * (last step's
// max) are evaluated exactly.  A cell left out has exact speed below
// SPEED_CAND_FRAC T / (1 - SPEED_EST_TOL) < T, so when the exact max of the
// evaluated cells is >= T it is the max over all cells; otherwise (the max
// fell by more than SPEED_DROP in one step) lf_speed_fallback rescans.
#define SPEED_DROP 1.0e-3
#define SPEED_CAND_FRAC 0.9999f
#define SPEED_EST_TOL 1.0e-5

__device__ __forceinline__ double speed_T(const StepArgs& a, int loop) {
    if (!loop) return 0.0;
    const double T = a.st->max_speed * (1.0 - SPEED_DROP);
    return (T > 1.0e-30) ? T : 0.0;  // also 0 for NaN
}

struct TileCtx {
    int m, l0, l1;
    uint32_t st_off;
    bool outlane, store;
    double dt, dz2, dz2i;
    float thr, gg1;
};

// One row of the march: output row l from the records S (l-1), C (l), N
// (l+1).  Marching down (rows l0 .. l1-1 in increasing order) the loads of
// row l+2 are issued first and land in S, which the caller rotates to the N
// slot; marching up (kUp: rows l1-1 .. l0) row l-2 lands in N.  The cell
// arithmetic is the same either way.
template <bool kSpeed, bool kBuf, class Div, bool kUp>
__device__ __forceinline__ void march_row(const StepArgs& a, const TileCtx& x, const StateIO<kBuf>& io, int l,
                                          Rec& S, const Rec& C, Rec& N, uint64_t& key, bool& rbad) {
    const int lp = kUp ? l - 2 : l + 2;
    const bool pre = kUp ? (lp >= x.l0 - 1) : (lp <= x.l1);
    double raw[8];
    if (pre) io.load_row(lp, raw);
    const LaneNb nb{C};
    const double dr2 = 2 * row_value(a.dr, l);
    double o[8];
    lf_cell_rad<Div>(N, S, C, nb, x.dt, x.dz2, x.dz2i, dr2, row_value(a.dr2inv, l), o);
    // With raw buffers every row issues the same eight stores, with no branch
    // around them: one counter tracks the wave's vector loads and stores,
    // and a path without them would make the compiler's wait for the
    // prefetched row (issued before the stores) wait for the stores too.
    const bool rb = (l == 1 || l == a.L);
    if (kBuf) io.store_out(l, rb, o, rb ? x.outlane : x.store);
    if (rb) {
        if (!kBuf && x.outlane) {
            char* rbp = reinterpret_cast<char*>(a.rowbuf + (int64_t)(l == 1 ? 0 : 8) * a.pitch) + x.st_off;
            for (int k = 0; k < 8; ++k) *reinterpret_cast<double*>(rbp + (int64_t)k * a.pitch * 8) = o[k];
        }
    } else {
        if (!kBuf && x.store) io.store_row(l, o);
        if (kSpeed && x.outlane) {
            const float est = speed_estimate(o, C.r, x.gg1);
            if (!(est < x.thr)) {
                const uint64_t kk = speed_key(wave_speed_u(o, C.r, a.gamma, a.gm1));
                key = (kk > key) ? kk : key;
            }
        }
    }
    // Rows 1 and L are screened by the boundary kernel, which writes them.
    if (x.outlane && !rb) rbad = rbad || !split_div_safe(o);
    if (pre) row_rec(a, lp, x.m, raw, kUp ? N : S);
}

// The production stencil.  Each wavefront owns strip s (columns 62 s ..
// 62 s + 61, one 512-byte slot per field row) for H rows: lane i holds
// column 62 s - 1 + i, so lanes 1..62 are the strip's cells and lanes 0 and
// 63 the neighbouring strips' edge cells, read in the same row load.  The
// wave marches down the rows with the records of rows l-1, l, l+1 in
// registers (three slots used round-robin, so nothing is copied between
// rows): every cell's conserved values are read from HBM once per tile and
// its physical values computed once; the m +- 1 neighbours come from the
// adjacent lanes by DPP shifts, the l +- 1 neighbours from the register
// window; each row's loads are issued a row ahead of their use.  The store of
// a row covers the whole slot: lane i writes slot (i + 63) mod 64, its own
// cell for lanes 1..62 and an unused double for lanes 0 and 63; lanes of the
// boundary and ghost columns 0, 1, M_max+1, M_max+2 do not store.
// One wavefront's tile (strip, chunk of H rows) of the march; `tile` and
// `up` (march the rows in decreasing order) are wave-uniform.
template <bool kSpeed, bool kBuf, class Div>
__device__ __forceinline__ void stencil_tile(const StepArgs& a, double dt_explicit, int loop, int H, int nstrips,
                                             int tile, bool up = false) {
    const int lane = (int)(threadIdx.x & 63);
    const int strip = tile % nstrips, chunk = tile / nstrips;
    TileCtx x;
    x.m = kStripCols * strip - 1 + lane;
    const int l0 = 1 + chunk * H;
    x.l0 = l0;
    x.l1 = min(l0 + H, a.L + 1);
    x.dt = step_dt(a, dt_explicit, loop);
    const bool central = x.m >= 2 && x.m <= a.Mn - 1;
    x.outlane = central && lane >= 1 && lane <= 62;
    x.store = x.outlane || lane == 0 || lane == 63 || x.m > a.Mn + 1;
    x.st_off = (uint32_t)(kSlotCols * strip + ((lane + 63) & 63)) * 8u;
    x.dz2 = a.dz2;
    x.dz2i = div_recip(a.dz2);
    x.thr = kSpeed ? (float)speed_T(a, loop) * SPEED_CAND_FRAC : 0.0f;
    x.gg1 = (float)(a.gamma * a.gm1);
    StateIO<kBuf> io;
    // Lanes outside the columns 0 .. M_max+2 read the nearest ghost column.
    io.init(a, (uint32_t)phys_col(min(max(x.m, 0), a.Mn + 1)) * 8u, x.st_off);
    Rec A, B, C;
    uint64_t key = 0;
    bool rbad = false;
    if (!up) {
        {
            double raw[8];
            io.load_row(l0 - 1, raw);
            row_rec(a, l0 - 1, x.m, raw, A);
            io.load_row(l0, raw);
            row_rec(a, l0, x.m, raw, B);
            io.load_row(l0 + 1, raw);
            row_rec(a, l0 + 1, x.m, raw, C);
        }
        int l = l0;
        for (; l + 3 <= x.l1; l += 3) {
            march_row<kSpeed, kBuf, Div, false>(a, x, io, l, A, B, C, key, rbad);
            march_row<kSpeed, kBuf, Div, false>(a, x, io, l + 1, B, C, A, key, rbad);
            march_row<kSpeed, kBuf, Div, false>(a, x, io, l + 2, C, A, B, key, rbad);
        }
        if (l < x.l1) march_row<kSpeed, kBuf, Div, false>(a, x, io, l, A, B, C, key, rbad);
        if (l + 1 < x.l1) march_row<kSpeed, kBuf, Div, false>(a, x, io, l + 1, B, C, A, key, rbad);
    } else {
        // Rows l1-1 down to l0: A, B, C start as rows l1-2, l1-1, l1 (all
        // within 0 .. L+1 since l0 >= 1 and l1 <= L+1).
        const int l1 = x.l1;
        {
            double raw[8];
            io.load_row(l1, raw);
            row_rec(a, l1, x.m, raw, C);
            io.load_row(l1 - 1, raw);
            row_rec(a, l1 - 1, x.m, raw, B);
            io.load_row(l1 - 2, raw);
            row_rec(a, l1 - 2, x.m, raw, A);
        }
        int l = l1 - 1;
        for (; l - 2 >= l0; l -= 3) {
            march_row<kSpeed, kBuf, Div, true>(a, x, io, l, A, B, C, key, rbad);
            march_row<kSpeed, kBuf, Div, true>(a, x, io, l - 1, C, A, B, key, rbad);
            march_row<kSpeed, kBuf, Div, true>(a, x, io, l - 2, B, C, A, key, rbad);
        }
        if (l >= l0) march_row<kSpeed, kBuf, Div, true>(a, x, io, l, A, B, C, key, rbad);
        if (l - 1 >= l0) march_row<kSpeed, kBuf, Div, true>(a, x, io, l - 1, C, A, B, key, rbad);
    }
    if (kSpeed) {
        key = wave_max_u64(key);
        if (lane == 0 && key != 0) atomicMax((unsigned long long*)&a.st->speed_bits, (unsigned long long)key);
    }
    flag_range(a.st, rbad);
}

template <bool kSpeed, bool kBuf, class Div>
__global__ void __launch_bounds__(256, 2) lf_stencil_tiled(StepArgs a, double dt_explicit, int loop, int H,
                                                           int nstrips, int ntiles, int order) {
    if (loop && loop_idle(a)) return;
    // Wave-uniform indices (scalar registers: row addressing and the march
    // loop then run on the scalar unit).  Order 0: tiles in launch order,
    // strip fastest: the four waves of a workgroup take four adjacent strips
    // of one band of rows, and the whole GPU works on a narrow band at a time
    // (an order giving each XCD its own contiguous band of rows measured 5%
    // slower).  Order 1: workgroups go to the XCDs round-robin (workgroup w
    // to XCD w mod 8), so workgroup w takes chunk 8 q + (w mod 8), strips
    // 4 g .. 4 g + 3 of it, with (q, g) counting the XCD's own workgroups:
    // each XCD sweeps whole chunks (all strips of a band of H rows), so the
    // edge columns two adjacent strips share are read into one L2, and the
    // XCDs still advance together over a narrow band of rows.  Order 2: the
    // four waves of a workgroup take four vertically adjacent tiles of one
    // strip (chunks 4 G .. 4 G + 3), waves 0 and 2 marching down, 1 and 3 up,
    // so each of the three rows two of them share (a tile's halo row is its
    // neighbour's first or last row) is read by both waves at the same point
    // of their marches, on one CU; the chunk groups go to the XCDs as the
    // chunks of order 1 (workgroup w: group 8 q + (w mod 8), all strips of it
    // one after another on that XCD).
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int tile;
    bool up = false;
    if (order == 0) {
        tile = (int)blockIdx.x * 4 + wave;
        if (tile >= ntiles) return;
    } else if (order == 3) {
        // Order 3: XCD x (workgroup w goes to XCD w mod 8) owns strips
        // s_lo(x) .. s_lo(x + 1) - 1, s_lo(x) = x * nstrips / 8, and takes
        // their tiles chunk by chunk: the two tiles a halo row is shared by
        // (the same strip, adjacent chunks) run on one XCD a few tiles
        // apart, so the second read of the row can hit that XCD's L2, while
        // the XCDs still advance together over a narrow band of rows.
        const int w = (int)blockIdx.x;
        const int x = w & 7, j = w >> 3;
        const int s0 = x * nstrips / 8, cnt = (x + 1) * nstrips / 8 - s0;
        if (cnt <= 0) return;
        const int i = 4 * j + wave;
        const int chunk = i / cnt;
        if (chunk >= ntiles / nstrips) return;
        tile = chunk * nstrips + s0 + (i - chunk * cnt);
    } else if (order == 2) {
        const int nchunks = ntiles / nstrips;
        const int w = (int)blockIdx.x;
        const int j = w >> 3;
        const int q = j / nstrips, strip = j - q * nstrips;
        const int chunk = 4 * ((q << 3) + (w & 7)) + wave;
        if (chunk >= nchunks) return;
        tile = chunk * nstrips + strip;
        up = (wave & 1) != 0;
    } else {
        const int gpc = (nstrips + 3) >> 2;  // workgroups per chunk
        const int w = (int)blockIdx.x;
        const int j = w >> 3;
        const int q = j / gpc, g = j - q * gpc;
        const int chunk = (q << 3) + (w & 7);
        const int strip = 4 * g + wave;
        if (chunk >= ntiles / nstrips || strip >= nstrips) return;
        tile = chunk * nstrips + strip;
    }
    stencil_tile<kSpeed, kBuf, Div>(a, dt_explicit, loop, H, nstrips, tile, up);
}

// The tiled stencil over many instances of uneven size in one launch.  The
// work list holds one entry per workgroup: its instance and the first of its
// four tiles in that instance (an instance's tiles are never shared with
// another instance's in one workgroup).  Workgroups stride over the list, so
// the grid may be smaller than the list (a persistent grid sized to the
// GPU).  Instance i's step reads args[2 i + (icur[i] ^ flip)]: its two
// ping-pong orientations, icur[i] being the state buffer it starts the graph
// from and flip the parity of the step within the graph.
template <bool kSpeed, bool kBuf, class Div>
__global__ void __launch_bounds__(256, 2) lf_stencil_tiled_batch(const StepArgs* __restrict__ args,
                                                                 const int* __restrict__ icur,
                                                                 const BatchBlock* __restrict__ work, int nwork,
                                                                 int flip, double dt_explicit, int loop) {
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int b = (int)blockIdx.x; b < nwork; b += (int)gridDim.x) {
        const BatchBlock w = work[b];
        const int inst = __builtin_amdgcn_readfirstlane(w.inst);
        const StepArgs a = args[2 * inst + (icur[inst] ^ flip)];
        if (loop && loop_idle(a)) continue;
        const int tile = __builtin_amdgcn_readfirstlane(w.first) + wave;
        if (tile >= w.count) continue;
        stencil_tile<kSpeed, kBuf, Div>(a, dt_explicit, loop, w.H, w.nstrips, tile);
    }
}

// ---------------------------------------------------- boundary conditions

// Per-thread work area in LDS for the boundary evaluation: the expression
// symbol table (read at run-time indices) and, for expressions deeper than
// the register stack, the evaluation stack.  Laid out [slot][lane] for the
// 64-thread workgroups of lf_bc_ring, so one slot of a wave is 64
// consecutive doubles.
#define BC_THREADS 64
static_assert(BC_THREADS == 64, "ring_lanes groups one kind of cell per 64-lane boundary workgroup");
struct WorkArea {
    double* base;  // LDS: bc_area_slots * BC_THREADS doubles (slot layout WA_* in lf_device.h)
    int lane;
    __device__ __forceinline__ double& at(int i) const { return base[i * BC_THREADS + lane]; }
    __device__ __forceinline__ double& sym(int i) const { return at(WA_SYM + i); }
    __device__ __forceinline__ double& stk(int i) const { return at(WA_STK + i); }
};

// The staged image a boundary kernel's StepArgs points to.
__device__ __forceinline__ const BcImage* img(const StepArgs& a) { return reinterpret_cast<const BcImage*>(a.bc); }

// The rarely used operations of the expression code, out of line so that
// the interpreter stays small where it is inlined.
__device__ __noinline__ double bc_fn1(int fn, double x) {
    switch (fn) {
        case MPD_FN_SIN: return sin(x);
        case MPD_FN_COS: return cos(x);
        case MPD_FN_TAN: return tan(x);
        case MPD_FN_ASIN: return asin(x);
        case MPD_FN_ACOS: return acos(x);
        case MPD_FN_ATAN: return atan(x);
        case MPD_FN_EXP: return exp(x);
        case MPD_FN_LOG: return log(x);
        case MPD_FN_LOG2: return log2(x);
        default: return log10(x);
    }
}

__device__ __noinline__ double bc_fn2(int op, double x, double y) {
    if (op == MPD_OP_MOD) return fmod(x, y);
    if (op == MPD_OP_POW) return pow(x, y);
    return atan2(x, y);
}

// Evaluate one compiled expression (the exprtk call sites of
// src/bcs/per_field_bc.cpp:323-358) on the symbol values in the work area,
// with the evaluation stack in the work area too: programs deeper than the
// register stack (out of line: one copy serves every call site).
__device__ __noinline__ double eval_expr_lds(const BcImage* bc, int off, int len, WorkArea w) {
    int sp = 0;
    for (int i = 0; i < len; ++i) {
        const MpdInstr in = bc->code(off + i);
        switch (in.op) {
            case MPD_OP_CONST: w.stk(sp++) = in.val; break;
            case MPD_OP_SYM: w.stk(sp++) = w.sym(in.arg); break;
            case MPD_OP_NEG: w.stk(sp - 1) = -w.stk(sp - 1); break;
            case MPD_OP_ADD: --sp; w.stk(sp - 1) = w.stk(sp - 1) + w.stk(sp); break;
            case MPD_OP_SUB: --sp; w.stk(sp - 1) = w.stk(sp - 1) - w.stk(sp); break;
            case MPD_OP_MUL: --sp; w.stk(sp - 1) = w.stk(sp - 1) * w.stk(sp); break;
            case MPD_OP_DIV: --sp; w.stk(sp - 1) = w.stk(sp - 1) / w.stk(sp); break;
            case MPD_OP_MOD:
            case MPD_OP_POW:
            case MPD_OP_ATAN2: --sp; w.stk(sp - 1) = bc_fn2(in.op, w.stk(sp - 1), w.stk(sp)); break;
            case MPD_OP_F1: {
                const double x = w.stk(sp - 1);
                double y;
                if (in.arg == MPD_FN_SQRT) y = sqrt(x);
                else if (in.arg == MPD_FN_ABS) y = fabs(x);
                else if (in.arg == MPD_FN_SIGN) y = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0);
                else y = bc_fn1(in.arg, x);
                w.stk(sp - 1) = y;
                break;
            }
            case MPD_OP_CLAMP: {
                const double lo = w.stk(sp - 3), x = w.stk(sp - 2), hi = w.stk(sp - 1);
                sp -= 2;
                w.stk(sp - 1) = (x < lo) ? lo : ((x > hi) ? hi : x);
                break;
            }
            case MPD_OP_MIN:
            case MPD_OP_MAX: {
                const int n = in.arg;
                double acc = w.stk(sp - n);
                for (int k = 1; k < n; ++k) {
                    const double v = w.stk(sp - n + k);
                    if (in.op == MPD_OP_MIN) acc = (v < acc) ? v : acc;
                    else acc = (acc < v) ? v : acc;
                }
                sp -= n - 1;
                w.stk(sp - 1) = acc;
                break;
            }
            default: break;
        }
    }
    return w.stk(0);
}

// Push and pop of the register stack s[0..kRegStack) (s[0] the top).
__device__ __forceinline__ void rs_push(double* s, double v) {
#pragma unroll
    for (int k = kRegStack - 1; k > 0; --k) s[k] = s[k - 1];
    s[0] = v;
}
// Replace the top n + 1 entries by r.
__device__ __forceinline__ void rs_reduce(double* s, int n, double r) {
    s[0] = r;
#pragma unroll
    for (int k = 1; k + n < kRegStack; ++k) s[k] = s[k + n];
}
__device__ __forceinline__ double rs_pick(bool is_min, double acc, double v) {
    return is_min ? ((v < acc) ? v : acc) : ((acc < v) ? v : acc);
}

// The same evaluation with the stack in registers, for programs whose
// deepest expression needs at most kRegStack entries.  The same operations
// in the same order as eval_expr_lds, so both give the same bits; the next
// instruction is read while the current one executes.
__device__ __forceinline__ double eval_expr_regs(const BcImage* bc, int off, int len, const WorkArea& w) {
    double s[kRegStack];
#pragma unroll
    for (int k = 0; k < kRegStack; ++k) s[k] = 0.0;
    if (len <= 0) return 0.0;
    MpdInstr next = bc->code(off);
    for (int i = 0; i < len; ++i) {
        const MpdInstr in = next;
        if (i + 1 < len) next = bc->code(off + i + 1);
        switch (in.op) {
            case MPD_OP_CONST: rs_push(s, in.val); break;
            case MPD_OP_SYM: rs_push(s, w.sym(in.arg)); break;
            case MPD_OP_NEG: s[0] = -s[0]; break;
            case MPD_OP_ADD: rs_reduce(s, 1, s[1] + s[0]); break;
            case MPD_OP_SUB: rs_reduce(s, 1, s[1] - s[0]); break;
            case MPD_OP_MUL: rs_reduce(s, 1, s[1] * s[0]); break;
            case MPD_OP_DIV: rs_reduce(s, 1, s[1] / s[0]); break;
            case MPD_OP_MOD:
            case MPD_OP_POW:
            case MPD_OP_ATAN2: rs_reduce(s, 1, bc_fn2(in.op, s[1], s[0])); break;
            case MPD_OP_F1: {
                const double x = s[0];
                double y;
                if (in.arg == MPD_FN_SQRT) y = sqrt(x);
                else if (in.arg == MPD_FN_ABS) y = fabs(x);
                else if (in.arg == MPD_FN_SIGN) y = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0);
                else y = bc_fn1(in.arg, x);
                s[0] = y;
                break;
            }
            case MPD_OP_CLAMP: {
                const double lo = s[2], x = s[1], hi = s[0];
                rs_reduce(s, 2, (x < lo) ? lo : ((x > hi) ? hi : x));
                break;
            }
            case MPD_OP_MIN:
            case MPD_OP_MAX: {
                // Arguments s[n-1] (first) .. s[0] (last), folded first to last.
                const bool mn = in.op == MPD_OP_MIN;
                if (in.arg == 2) {
                    rs_reduce(s, 1, rs_pick(mn, s[1], s[0]));
                } else if (in.arg == 3) {
                    rs_reduce(s, 2, rs_pick(mn, rs_pick(mn, s[2], s[1]), s[0]));
                } else if (in.arg == 4) {
                    rs_reduce(s, 3, rs_pick(mn, rs_pick(mn, rs_pick(mn, s[3], s[2]), s[1]), s[0]));
                }
                break;
            }
            default: break;
        }
    }
    return s[0];
}
static_assert(kRegStack == 4, "eval_expr_regs folds MIN/MAX of up to four arguments");

// Physical slot of each condition (evaluation order rho, v_z, v_r, v_phi,
// H_z, H_r, H_phi, e) and of each neighbour symbol (SetNeighbors order rho,
// v_z, v_r, v_phi, e, H_z, H_r, H_phi).
// (Functions, not tables: indexed by unrolled loop counters they fold to
// constants and the state arrays stay in registers.)
__device__ __forceinline__ constexpr int cond_slot(int k) {
    return k == 0 ? PH_RHO : k == 1 ? PH_VZ : k == 2 ? PH_VR : k == 3 ? PH_VPHI
         : k == 4 ? PH_HZ : k == 5 ? PH_HR : k == 6 ? PH_HPHI : PH_E;
}
__device__ __forceinline__ constexpr int nb_slot(int k) {
    return k == 0 ? PH_RHO : k == 1 ? PH_VZ : k == 2 ? PH_VR : k == 3 ? PH_VPHI
         : k == 4 ? PH_E : k == 5 ? PH_HZ : k == 6 ? PH_HR : PH_HPHI;
}

__device__ __forceinline__ bool seg_covers(const MpdSeg& s, int i) { return i >= s.lo && i <= s.hi; }

__device__ __forceinline__ bool face_covers(const BcImage* bc, int f, int i) {
    for (int s = 0; s < bc->nseg[f]; ++s)
        if (seg_covers(bc->seg(f, s), i)) return true;
    return false;
}

// The fixed-axis index of a face and its interior neighbour
// (src/bcs/per_field_bc.cpp:264-274).
__device__ __forceinline__ void face_nb(const StepArgs& a, int f, int l, int m, int& ln, int& mn) {
    ln = l;
    mn = m;
    if (f == MPD_L_LO) ln = 2;
    else if (f == MPD_L_HI) ln = a.L - 1;
    else if (f == MPD_M_LO) mn = 2;
    else mn = a.Mn - 1;
}

// Geometry symbols of one face-chain evaluation at a cell: its row's dr,
// its r, r_z and z, and its face neighbour's r and r_z.
struct ChainGeo {
    double dr, r, rz, z, rnb, rznb;
};

// The same read from the per-row wall data (the AxisLF stencil's l +- 1
// chains, outside the boundary geometry table's cells).
__device__ __forceinline__ ChainGeo chain_geo_rows(const StepArgs& a, int f, int l, int m) {
    int ln, mn;
    face_nb(a, f, l, m, ln, mn);
    return ChainGeo{a.dr[l], cell_r(a, l, m), cell_rz(a, l, m), (double)(l - 1) * a.dz, cell_r(a, ln, mn),
                    cell_rz(a, ln, mn)};
}

// The boundary geometry table's entries of cell (l, m) (lf_device.h).
struct RingGeo {
    double v[RG_N];
};
__device__ __forceinline__ void load_ring_geo(const StepArgs& a, int l, int m, RingGeo& g) {
    const int64_t n = ring_lanes_compact(a.L, a.Mn);
    const int c = ring_compact_index(l, m, a.L, a.Mn);
#pragma unroll
    for (int e = 0; e < RG_N; ++e) g.v[e] = a.ringgeo[e * n + c];
}

// Apply the segments [0, seg_end) of face f that cover cell (l, m) to the
// physical state cur[10], with the neighbour state nbr[10] (both indexed by
// PH_*): PerFieldBC::Apply's per-cell physical part
// (src/bcs/per_field_bc.cpp:297-360, 391-446).  Returns the index of the
// last segment applied, or -1.
// Per segment the Neumann and Dirichlet conditions resolve first, in
// registers: they read only the neighbour state and constants.  The
// expressions then run in field order; an expression reads only the fields
// evaluated before it (forward references are rejected when the program is
// built), so each sees the values the serial order rho, v_z, ..., e gives it.
__device__ __forceinline__ int face_chain(const StepArgs& a, const WorkArea& w, int f, int l, int m, int seg_end,
                                          const ChainGeo& cg, double* cur, const double* nbr) {
    const BcImage* bc = img(a);
    const bool is_l = (f == MPD_L_LO || f == MPD_L_HI);
    const int i = is_l ? m : l;
    const bool regs = a.bc_area - WA_STK <= kRegStack;
    int last = -1;
    for (int s = 0; s < seg_end; ++s) {
        const MpdSeg& g = bc->seg(f, s);
        if (!seg_covers(g, i)) continue;
        last = s;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const MpdCond& c = g.c[k];
            const int slot = cond_slot(k);
            if (c.type == MPD_NEUMANN) cur[slot] = nbr[slot] + c.value;
            else if (c.type == MPD_DIRICHLET) cur[slot] = c.value;
        }
        if (!g.has_expr) continue;
        w.sym(MPD_BS_GAMMA) = a.gamma;
        w.sym(MPD_BS_BETA) = a.beta;
        w.sym(MPD_BS_HZ0) = a.H_z0;
        w.sym(MPD_BS_R0) = a.r_0;
        w.sym(MPD_BS_DZ) = a.dz;
        w.sym(MPD_BS_DR) = cg.dr;
        w.sym(MPD_BS_R) = cg.r;
        w.sym(MPD_BS_RZ) = cg.rz;
        w.sym(MPD_BS_Z) = cg.z;
        w.sym(MPD_BS_RNB) = cg.rnb;
        w.sym(MPD_BS_RZNB) = cg.rznb;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            w.sym(MPD_BS_RHO_NB + k) = nbr[nb_slot(k)];
            w.sym(MPD_BS_RHO + k) = cur[cond_slot(k)];
        }
#pragma unroll 1
        for (int k = 0; k < 8; ++k) {
            const MpdCond& c = g.c[k];
            if (c.type != MPD_EXPRESSION) continue;
            w.sym(MPD_BS_RHO + k) = regs ? eval_expr_regs(bc, c.code_off, c.code_len, w)
                                         : eval_expr_lds(bc, c.code_off, c.code_len, w);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (g.c[k].type == MPD_EXPRESSION) cur[cond_slot(k)] = w.sym(MPD_BS_RHO + k);
    }
    return last;
}

// RebuildUFromPhysical, include/physics_utils.hpp:15-24.
__device__ __forceinline__ void rebuild_u(const double* ph, double r, double* u) {
    u[0] = ph[PH_RHO] * r;
    u[1] = ph[PH_RHO] * ph[PH_VZ] * r;
    u[2] = ph[PH_RHO] * ph[PH_VR] * r;
    u[3] = ph[PH_RHO] * ph[PH_VPHI] * r;
    u[4] = ph[PH_RHO] * ph[PH_E] * r;
    u[5] = ph[PH_HPHI];
    u[6] = ph[PH_HZ] * r;
    u[7] = ph[PH_HR] * r;
}

// State of boundary cell (l, 1) as the M_LO face sees it while applying its
// segment s (serial order, l ascending): the M_LO segments before s, plus
// segment s itself when `with_s` (the cell was already visited in s).  Cell
// (1,1) belongs to L_LO, which ran before; cell (L,1) to L_HI, which runs
// after.  In two parts: mlo_plan issues the loads, mlo_finish runs the chain.
struct MloPlan {
    CellSrc s, n;
    int l, cf, ce, nl, nm;
    double v[10], vn[10];
};

__device__ __forceinline__ void mlo_plan(const StepArgs& a, int l, int s, bool with_s, MloPlan& p) {
    p.l = l;
    p.cf = -1;
    p.ce = 0;
    p.nl = l;
    p.nm = 2;
    if (l == 1) {
        if (face_covers(img(a), MPD_L_LO, 1)) {
            p.cf = MPD_L_LO;
            p.ce = img(a)->nseg[MPD_L_LO];
            p.nl = 2;
            p.nm = 1;
        }
    } else if (l != a.L) {
        p.cf = MPD_M_LO;
        p.ce = with_s ? s + 1 : s;
    }
    p.s = s0_src(a, l, 1);
    p.n = s0_src(a, p.nl, p.nm);
    src_load(p.s, p.v);
    if (p.cf >= 0) src_load(p.n, p.vn);
}

__device__ __forceinline__ void mlo_finish(const StepArgs& a, const WorkArea& w, const MloPlan& p, double* ph) {
    src_phys(p.s, p.v, cell_r(a, p.l, 1), a.gm1, ph);
    if (p.cf >= 0) {
        double nb[10];
        src_phys(p.n, p.vn, cell_r(a, p.nl, p.nm), a.gm1, nb);
        face_chain(a, w, p.cf, p.l, 1, p.ce, chain_geo_rows(a, p.cf, p.l, 1), ph, nb);
    }
}

// PerFieldBC::AxisLfU1/U2/U5/U7, src/bcs/per_field_bc.cpp:490-566, at (l, 1).
// own: the cell's physical values after the segment; lo/hi: the states of
// (l-1, 1) and (l+1, 1); c2: the state of (l, 2); sp: start-of-step physical
// values at (l-1,1), (l,1), (l+1,1) (p and P are never written by the BCs).
__device__ __forceinline__ void axis_overrides(const StepArgs& a, const MpdSeg& g, int l, double dt, const double* own,
                               const double* lo, const double* hi, const double* c2, const double* sp_lo,
                               const double* sp_c, const double* sp_hi, double* u) {
    const int m = 1;
    const double r_n = cell_r(a, l + 1, m), r_s = cell_r(a, l - 1, m);
    const double r_e = cell_r(a, l, m + 1), r_c = cell_r(a, l, m);
    const double dr_l = a.dr[l], dz = a.dz;
    auto U0 = [&](int k, int ll, int mm) { return fld(a.u_in, a.fstride, k)[at(a, ll, mm)]; };
    if (g.c[0].type == MPD_AXIS_LF) {
        const int k = 0;
        const double n = U0(k, l + 1, m), s = U0(k, l - 1, m), e = U0(k, l, m + 1), c = U0(k, l, m);
        u[0] = (0.25 * (n / r_n + s / r_s + e / r_e + c / r_c) +
                dt * (-(n / r_n * hi[PH_VZ] - s / r_s * lo[PH_VZ]) / (2.0 * dz) -
                      (e / r_e * c2[PH_VR] - c / r_c * c2[PH_VR]) / dr_l)) *
               r_c;
    }
    if (g.c[1].type == MPD_AXIS_LF) {
        const int k = 1;
        const double n = U0(k, l + 1, m), s = U0(k, l - 1, m), e = U0(k, l, m + 1), c = U0(k, l, m);
        u[1] = (0.25 * (n / r_n + s / r_s + e / r_e + c / r_c) +
                dt * (((hi[PH_HZ] * hi[PH_HZ] - sp_hi[PH_PT]) - (lo[PH_HZ] * lo[PH_HZ] - sp_lo[PH_PT])) / (2.0 * dz) +
                      (c2[PH_HZ] * c2[PH_HR] - own[PH_HZ] * own[PH_HR]) / dr_l -
                      (n / r_n * hi[PH_VZ] - s / r_s * lo[PH_VZ]) / (2.0 * dz) -
                      (e / r_e * c2[PH_VR] - c / r_c * own[PH_VR]) / dr_l)) *
               r_c;
    }
    if (g.c[7].type == MPD_AXIS_LF) {
        const int k = 4;
        const double n = U0(k, l + 1, m), s = U0(k, l - 1, m), e = U0(k, l, m + 1), c = U0(k, l, m);
        u[4] = (0.25 * (n / r_n + s / r_s + e / r_e + c / r_c) +
                dt * (-sp_c[PH_P] * ((hi[PH_VZ] - lo[PH_VZ]) / (2.0 * dz) + (c2[PH_VR] - own[PH_VR]) / dr_l) -
                      (n / r_n * hi[PH_VZ] - s / r_s * lo[PH_VZ]) / (2.0 * dz) -
                      (e / r_e * c2[PH_VR] - c / r_c * own[PH_VR]) / dr_l)) *
               r_c;
    }
    if (g.c[4].type == MPD_AXIS_LF) {
        const int k = 6;
        const double n = U0(k, l + 1, m), s = U0(k, l - 1, m), e = U0(k, l, m + 1), c = U0(k, l, m);
        u[6] = (0.25 * (n / r_n + s / r_s + e / r_e + c / r_c) +
                dt * ((c2[PH_HR] * c2[PH_VZ] - own[PH_HR] * own[PH_VZ]) / dr_l -
                      (e / r_e * c2[PH_VR] - c / r_c * own[PH_VR]) / dr_l)) *
               r_c;
    }
}

// Final conserved values of boundary cell (l, m) written by face f (the last
// face in L_LO, M_HI, M_LO, L_HI order that covers it); G: the cell's
// boundary geometry table entries.
__device__ __forceinline__ void bc_final(const StepArgs& a, const WorkArea& w, int f, int l, int m, double dt,
                                         double* u, const RingGeo& G, const double* pre_o, const double* pre_n,
                                         int pre_nl, int pre_nm) {
    // Neighbour of the face (src/face_bc.cpp order: L_LO, M_HI, M_LO, L_HI;
    // only corners see an earlier face at their neighbour).
    int ln, mn;
    face_nb(a, f, l, m, ln, mn);
    int pf = -1, pl = 0, pm = 0, ql = 0, qm = 0;  // earlier face at the neighbour, its cell and neighbour
    if (f == MPD_M_HI && ln == 1 && face_covers(img(a), MPD_L_LO, mn)) {
        pf = MPD_L_LO; pl = 1; pm = mn; ql = 2; qm = mn;
    } else if (f == MPD_L_HI && mn == 1 && face_covers(img(a), MPD_M_LO, ln)) {
        pf = MPD_M_LO; pl = ln; pm = 1; ql = ln; qm = 2;
    } else if (f == MPD_L_HI && mn == a.Mn && face_covers(img(a), MPD_M_HI, ln)) {
        pf = MPD_M_HI; pl = ln; pm = a.Mn; ql = ln; qm = a.Mn - 1;
    }
    // The own and positional-neighbour states were loaded with the lane
    // (ring_lane); the face neighbour where it is not the positional one, and
    // the earlier face's neighbour, here.
    const CellSrc so = s0_src(a, l, m), sn = s0_src(a, ln, mn);
    const CellSrc sq = s0_src(a, pf >= 0 ? ql : ln, pf >= 0 ? qm : mn);
    double vo[10], vn[10], vq[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) vo[k] = pre_o[k];
    if (ln == pre_nl && mn == pre_nm) {
#pragma unroll
        for (int k = 0; k < 10; ++k) vn[k] = pre_n[k];
    } else {
        src_load(sn, vn);
    }
    if (pf >= 0) src_load(sq, vq);
    // The face neighbour is the row neighbour for an L face, the column
    // neighbour for an M face; an earlier face's chain at a corner's
    // neighbour reads the diagonal cell (table entries, lf_device.h).
    const bool lf = (f == MPD_L_LO || f == MPD_L_HI);
    const double r = G.v[RG_R];
    const double r_nb = lf ? G.v[RG_R_ROW] : G.v[RG_R_COL], rz_nb = lf ? G.v[RG_RZ_ROW] : G.v[RG_RZ_COL];
    double st[10], nb[10];
    src_phys(so, vo, r, a.gm1, st);
    src_phys(sn, vn, r_nb, a.gm1, nb);
    if (pf >= 0) {
        double nb2[10];
        src_phys(sq, vq, G.v[RG_R_DIAG], a.gm1, nb2);
        // M_HI's corner (1, Mn): L_LO's chain at (1, Mn - 1), the column
        // neighbour; L_HI's corners: M_LO's / M_HI's at (L - 1, m), the row
        // neighbour.  Either way the chain's own neighbour is the diagonal.
        const ChainGeo pg = (f == MPD_M_HI)
            ? ChainGeo{G.v[RG_DR], G.v[RG_R_COL], G.v[RG_RZ_COL], G.v[RG_Z], G.v[RG_R_DIAG], G.v[RG_RZ_DIAG]}
            : ChainGeo{G.v[RG_DR_ROW], G.v[RG_R_ROW], G.v[RG_RZ_ROW], G.v[RG_Z_ROW], G.v[RG_R_DIAG],
                       G.v[RG_RZ_DIAG]};
        face_chain(a, w, pf, pl, pm, img(a)->nseg[pf], pg, nb, nb2);
    }
    const int last = face_chain(a, w, f, l, m, img(a)->nseg[f],
                                ChainGeo{G.v[RG_DR], r, G.v[RG_RZ], G.v[RG_Z], r_nb, rz_nb}, st, nb);
    rebuild_u(st, r, u);
    if (f == MPD_M_LO && last >= 0 && img(a)->seg(f, last).has_axis) {
        const MpdSeg& g = img(a)->seg(f, last);
        MloPlan plo, phi;
        mlo_plan(a, l - 1, last, seg_covers(g, l - 1), plo);
        mlo_plan(a, l + 1, last, false, phi);
        // The start-of-step values at (l-1,1), (l,1), (l+1,1) are the
        // values loaded for the plans' own states and for this cell's.
        double lo[10], hi[10], sp_lo[10], sp_c[10], sp_hi[10];
        src_phys(plo.s, plo.v, cell_r(a, l - 1, 1), a.gm1, sp_lo);
        src_phys(phi.s, phi.v, cell_r(a, l + 1, 1), a.gm1, sp_hi);
        src_phys(so, vo, r, a.gm1, sp_c);
        mlo_finish(a, w, plo, lo);
        mlo_finish(a, w, phi, hi);
        axis_overrides(a, g, l, dt, st, lo, hi, nb, sp_lo, sp_c, sp_hi, u);
    }
}

// The loop-state fields the end of a step reads, all loaded before any is
// written (one round trip instead of a chain of dependent ones).
struct LoopRead {
    double t, dt_use, max_speed;
    uint64_t bits;
    int32_t step, max_steps, idle;
    uint32_t range;
};

__device__ __forceinline__ LoopRead read_loop(const MpdLoopState* st) {
    LoopRead r;
    r.bits = __hip_atomic_load(&st->speed_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r.range = __hip_atomic_load(&st->range_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r.t = st->t;
    r.dt_use = st->dt_use;
    r.max_speed = st->max_speed;
    r.step = st->step;
    r.max_steps = st->max_steps;
    r.idle = st->done | st->fallback | st->range_stop;
    return r;
}

// dt bookkeeping of the time loop after one step: t += dt; dt =
// ComputeDt(...) when adaptive (src/diagnostics.cpp:49-65); the next step's
// dt truncated to T (src/main.cpp:220-223).
__device__ void finalize_from(const StepArgs& a, const LoopRead& r) {
    MpdLoopState* st = a.st;
    const double speed = __longlong_as_double((long long)r.bits);
    const double dt_used = r.dt_use;
    const double t = r.t + dt_used;
    const int step = r.step + 1;
    double dt = dt_used;
    if (a.adaptive) {
        const double dt_cfl = a.cfl_dx / (speed + 1.0e-10);
        const double dt_grown = dt_used * a.growth;
        const double v = (dt_grown < dt_cfl) ? dt_grown : dt_cfl;
        dt = (v < a.dt_min) ? a.dt_min : ((a.dt_max < v) ? a.dt_max : v);
    }
    const bool cap = r.max_steps > 0 && step >= r.max_steps;
    st->max_speed = speed;
    st->t = t;
    st->step = step;
    st->dt = dt;
    st->dt_use = (t + dt > a.T) ? a.T - t : dt;
    st->done = (!(t < a.T) || cap) ? 1 : 0;
    st->speed_bits = 0;
    st->ticket = 0;
    // A new state outside the split division's range: the loop idles after
    // this (complete) step until the host switches to IEEE division.
    if (r.range != 0) st->range_stop = 1;
}

__device__ void finalize_step(const StepArgs& a) { finalize_from(a, read_loop(a.st)); }

// Copies the program image (words 8-byte words) into `prog` (LDS).  Every
// thread issues all its loads before its LDS stores: one global round trip
// per 512 words (a plain strided copy loop waits for each load in turn).
__device__ __forceinline__ void stage_image(const uint64_t* src, uint64_t* prog, int words) {
    constexpr int kPer = 8;
    for (int base = 0; base < words; base += kPer * BC_THREADS) {
        uint64_t v[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = base + j * BC_THREADS + (int)threadIdx.x;
            v[j] = (i < words) ? src[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = base + j * BC_THREADS + (int)threadIdx.x;
            if (i < words) prog[i] = v[j];
        }
    }
    __syncthreads();
}

// One thread per boundary cell: row 1, row L, column 1 (l = 2..L-1), column
// Mn (l = 2..L-1).  Writes each cell's final conserved values and adds its
// wave speed to the running max; prog and area are the block's LDS.  In the
// time loop the step is finished after the boundary kernel by a kernel of
// its own (lf_finish_one, lf_finish_batch): the kernel boundary orders every
// wave-speed atomic of the step before it, so no workgroup waits for a
// completion ticket or a device-scope fence (measured 25.5 -> 21.5 us per
// boundary launch at 6400x2560).
// A boundary lane's cell and its geometry table entries, read before the
// program image is staged (the loads share the staging's round trip).
// A boundary lane's cell, its geometry table entries, and the start states of
// the cell and of its face neighbour by position (row 1: (2, m), row L:
// (L-1, m), column 1: (l, 2), column Mn: (l, Mn-1)), loaded when the lane is
// set up, before the program image is staged, so that the loads overlap the
// staging.  The face neighbour differs from the positional one only at the
// corners (1, Mn) and (L, Mn) when M_HI owns them; bc_final loads it there.
struct RingLane {
    bool valid;
    int l, m;
    int nl, nm;
    RingGeo G;
    double vo[10], vn[10];
};

__device__ __forceinline__ RingLane ring_lane(const StepArgs& a, int idx, bool grouped) {
    RingLane r;
    r.l = r.m = 0;
    r.valid = grouped ? ring_cell(idx, a.L, a.Mn, r.l, r.m) : ring_cell_compact(idx, a.L, a.Mn, r.l, r.m);
    if (r.l == 1) {
        r.nl = 2;
        r.nm = r.m;
    } else if (r.l == a.L) {
        r.nl = a.L - 1;
        r.nm = r.m;
    } else if (r.m == 1) {
        r.nl = r.l;
        r.nm = 2;
    } else {
        r.nl = r.l;
        r.nm = a.Mn - 1;
    }
    if (r.valid) {
        load_ring_geo(a, r.l, r.m, r.G);
        src_load(s0_src(a, r.l, r.m), r.vo);
        src_load(s0_src(a, r.nl, r.nm), r.vn);
    }
    return r;
}

__device__ __forceinline__ void bc_ring_block(const StepArgs& a_in, double dt, int speed, const uint64_t* prog,
                                              double* area, const RingLane& rl) {
    StepArgs a = a_in;
    a.bc = prog;  // the staged image, read through img(a)
    const WorkArea w{area, (int)(threadIdx.x & (BC_THREADS - 1))};
    const int L = a.L, Mn = a.Mn;
    uint64_t key = 0;
    bool rbad = false;
    const int l = rl.l, m = rl.m;
    if (rl.valid) {
        const RingGeo& G = rl.G;
        // Last face covering the cell (M_LO skips l = 1 and l = L by its corner
        // policy, so it never owns a row-1 or row-L cell).
        int f = -1;
        const BcImage* bc = img(a);
        if (l == 1) {
            if (m == Mn && face_covers(bc, MPD_M_HI, 1)) f = MPD_M_HI;
            else if (face_covers(bc, MPD_L_LO, m)) f = MPD_L_LO;
        } else if (l == L) {
            if (face_covers(bc, MPD_L_HI, m)) f = MPD_L_HI;
            else if (m == Mn && face_covers(bc, MPD_M_HI, L)) f = MPD_M_HI;
        } else if (m == 1) {
            if (face_covers(bc, MPD_M_LO, l)) f = MPD_M_LO;
        } else {
            if (face_covers(bc, MPD_M_HI, l)) f = MPD_M_HI;
        }
        double u[8];
        if (f >= 0) {
            bc_final(a, w, f, l, m, dt, u, G, rl.vo, rl.vn, rl.nl, rl.nm);
        } else if (m >= 2 && m <= Mn - 1) {
            // Uncovered row cell keeps its central value.
            const double* rb = a.rowbuf + (int64_t)(l == 1 ? 0 : 8) * a.pitch;
            for (int k = 0; k < 8; ++k) u[k] = rb[(int64_t)k * a.pitch + phys_col(m)];
        } else {
            // Uncovered column cell: the reference never writes u there, so u
            // (and u0 after the copy) stays 0 whatever the start state held.
            for (int k = 0; k < 8; ++k) u[k] = 0.0;
        }
        {
            const int64_t i = at(a, l, m);
            for (int k = 0; k < 8; ++k) a.u_out[(int64_t)k * a.fstride + i] = u[k];
        }
        if (speed) key = speed_key(wave_speed_u(u, G.v[RG_R], a.gamma, a.gm1));
        // An uncovered column cell holds u = 0 (never read by a division
        // with a nonzero numerator: its u_1 is 0, the state has diverged).
        rbad = (f >= 0 || (m >= 2 && m <= Mn - 1)) && !split_div_safe(u);
    }
    if (speed) {
        key = wave_max_u64(key);
        if ((threadIdx.x & 63) == 0 && key != 0) atomicMax((unsigned long long*)&a.st->speed_bits, (unsigned long long)key);
    }
    flag_range(a.st, rbad);
}

__global__ void __launch_bounds__(BC_THREADS) lf_bc_ring(StepArgs a_in, double dt_explicit, int loop, int speed) {
    // The BC program (segments and the expression code in use) is read by
    // every thread many times over in dependent chains: its image goes to
    // LDS once, the per-lane work area behind it.  The loop-state reads and
    // the image loads are issued together (one round trip), the idle test
    // taken after.
    extern __shared__ uint64_t bc_lds[];
    const bool idle = loop && loop_idle(a_in);
    const double dt = step_dt(a_in, dt_explicit, loop);
    const RingLane rl = ring_lane(a_in, (int)blockIdx.x * BC_THREADS + (int)threadIdx.x, true);
    stage_image(a_in.bc, bc_lds, a_in.bc_words);
    if (idle) return;
    bc_ring_block(a_in, dt, speed, bc_lds, reinterpret_cast<double*>(bc_lds + a_in.bc_words), rl);
}

// The boundary kernel over many instances: work entry b = (instance, block
// index within the instance's ring, the instance's block count);
// lf_finish_batch finishes the instances' steps after it.
__global__ void __launch_bounds__(BC_THREADS) lf_bc_ring_batch(const StepArgs* __restrict__ args,
                                                               const int* __restrict__ icur,
                                                               const BatchBlock* __restrict__ work, int flip,
                                                               int speed, int img_words) {
    const BatchBlock w = work[blockIdx.x];
    const StepArgs a = args[2 * w.inst + (icur[w.inst] ^ flip)];
    // img_words: the largest image of the launch's instances (the work area's
    // offset); this instance's image is a.bc_words long.
    extern __shared__ uint64_t bc_lds[];
    const bool idle = loop_idle(a);
    const double dt = step_dt(a, 0.0, 1);
    const RingLane rl = ring_lane(a, w.first * BC_THREADS + (int)threadIdx.x, false);
    stage_image(a.bc, bc_lds, a.bc_words);
    if (idle) return;
    bc_ring_block(a, dt, speed, bc_lds, reinterpret_cast<double*>(bc_lds + img_words), rl);
}

// The step's end for every instance of a batch that took it (after the
// boundary kernel, so every wave-speed contribution of the step is in): the
// exact-rescan test and the dt bookkeeping of lf_bc_ring's last block, one
// thread per instance.  The kernel boundary orders it after all of the step's
// atomics, so no per-workgroup ticket or device-scope fence is needed.
// The end of one instance's step once all its cells are written: the
// exact-rescan test, then the dt bookkeeping (finalize_step).
// r: the loop state read at the step's end (read_loop).
__device__ __forceinline__ void finish_instance(const StepArgs& a, const LoopRead& r) {
    // speed_T with the loop state already read.
    const double Tr = r.max_speed * (1.0 - SPEED_DROP);
    const double T = (Tr > 1.0e-30) ? Tr : 0.0;
    const double mc = __longlong_as_double((long long)r.bits);
    if (T > 0.0 && !(mc >= T)) {
        a.st->fallback = 1;
        a.st->ticket = 0;
    } else {
        finalize_from(a, r);
    }
}

// The step's end of a single grid's loop step, after lf_bc_ring.
__global__ void lf_finish_one(StepArgs a) {
    const LoopRead r = read_loop(a.st);
    if (r.idle) return;
    finish_instance(a, r);
}

__global__ void __launch_bounds__(64) lf_finish_batch(const StepArgs* __restrict__ args, const int* __restrict__ icur,
                                                      int n, int flip) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const StepArgs a = args[2 * i + (icur[i] ^ flip)];
    const LoopRead r = read_loop(a.st);
    if (r.idle) return;
    finish_instance(a, r);
}

// Exact max wave speed over the whole new state, run when the candidate max
// of the step could not be proven to be the max (see SPEED_DROP); then the
// step's dt bookkeeping.  Exits at once otherwise.
__global__ void __launch_bounds__(256) lf_speed_fallback(StepArgs a) {
    if (!a.st->fallback) return;
    const int64_t n = (int64_t)a.L * a.Mn;
    uint64_t key = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int l = 1 + (int)(c / a.Mn), m = 1 + (int)(c % a.Mn);
        double u[8];
        for (int k = 0; k < 8; ++k) u[k] = a.u_out[(int64_t)k * a.fstride + at(a, l, m)];
        const uint64_t kk = speed_key(wave_speed_u(u, cell_r(a, l, m), a.gamma, a.gm1));
        key = (kk > key) ? kk : key;
    }
    key = wave_max_u64(key);
    if ((threadIdx.x & 63) == 0 && key != 0) atomicMax((unsigned long long*)&a.st->speed_bits, (unsigned long long)key);
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned t = atomicAdd(&a.st->ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        a.st->fallback = 0;
        finalize_step(a);
    }
}

// -------------------------------------------------------------- reductions

// SolutionChange over the 7 fields rho, v_z, v_r, v_phi, H_z, H_r, H_phi of the
// interior, then SavePrev when `save`.  prev: 7 fields, same layout.
// partial: 2 doubles per block, summed in block order by the last block.
// In the loop (`loop`), it runs only on a step that is a multiple of
// check_frequency and was not checked yet, and sets converged/done
// (src/main.cpp:233-247).

// The sums of one 256-thread block's share of SolutionChange (and the
// snapshot update when `save`): block `bid` of `nblk` strides over the
// interior cells; the squares are summed per thread, then over the block by a
// fixed tree.  The block's two sums are returned in thread 0.
__device__ __forceinline__ void residual_sums(const StepArgs& a, double* prev, int save, int bid, int nblk,
                                              double& bd, double& bc) {
    const int L = a.L, Mn = a.Mn;
    const int64_t n = (int64_t)L * Mn;
    double sd = 0.0, sc = 0.0;
    for (int64_t c = (int64_t)bid * blockDim.x + threadIdx.x; c < n; c += (int64_t)nblk * blockDim.x) {
        const int l = 1 + (int)(c / Mn), m = 1 + (int)(c % Mn);
        const int64_t i = at(a, l, m);
        double ph[10];
        start_phys(a, l, m, ph);
        const int ids[7] = {PH_RHO, PH_VZ, PH_VR, PH_VPHI, PH_HZ, PH_HR, PH_HPHI};
        double dsum = 0.0, csum = 0.0;
        for (int k = 0; k < 7; ++k) {
            const double cur = ph[ids[k]];
            const double d = cur - prev[(int64_t)k * a.fstride + i];
            dsum += d * d;
            csum += cur * cur;
            if (save) prev[(int64_t)k * a.fstride + i] = cur;
        }
        sd += dsum;
        sc += csum;
    }
    __shared__ double sh_d[256], sh_c[256];
    __syncthreads();  // the arrays' previous use (a caller looping over blocks) is complete
    sh_d[threadIdx.x] = sd;
    sh_c[threadIdx.x] = sc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh_d[threadIdx.x] += sh_d[threadIdx.x + s];
            sh_c[threadIdx.x] += sh_c[threadIdx.x + s];
        }
        __syncthreads();
    }
    bd = sh_d[0];
    bc = sh_c[0];
}

// SolutionChange from the sums over all blocks (gd: squared differences, gc:
// squared values), and the loop's convergence stop (src/main.cpp:233-247).
__device__ __forceinline__ void residual_result(const StepArgs& a, double gd, double gc, int loop, double threshold) {
    const double nd = sqrt(gd), nc = sqrt(gc);
    const double change = (nc > 1e-15) ? (nd / nc) : nd;
    a.st->change = change;
    if (loop) {
        a.st->checked_step = a.st->step;
        if (change < threshold) {
            a.st->converged = 1;
            a.st->done = 1;
        }
    }
}

// Diagnostics::SolutionChange + Fields::SavePrev (src/diagnostics.cpp:71-106,
// src/fields.cpp:59-74) as block `bid` of `nblk`: the last block to finish
// adds the block sums in block order.
__device__ __forceinline__ void residual_block(const StepArgs& a, double* prev, double* partial, int loop,
                                               int check_frequency, double threshold, int save, int bid, int nblk) {
    double bd, bc;
    residual_sums(a, prev, save, bid, nblk, bd, bc);
    __shared__ bool last;
    if (threadIdx.x == 0) {
        partial[2 * bid] = bd;
        partial[2 * bid + 1] = bc;
        __threadfence();
        const unsigned t = atomicAdd(&a.st->ticket2, 1u);
        last = (t == (unsigned)nblk - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double gd = 0.0, gc = 0.0;
        for (int b = 0; b < nblk; ++b) {
            gd += __hip_atomic_load(&partial[2 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            gc += __hip_atomic_load(&partial[2 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        residual_result(a, gd, gc, loop, threshold);
        a.st->ticket2 = 0;
    }
}

// Whether the loop checks convergence after the step just finished: the step
// count is a multiple of check_frequency, no rescan is pending,
// and was not checked yet (src/main.cpp:233-247).
__device__ __forceinline__ bool residual_due(const StepArgs& a, int check_frequency) {
    const int step = a.st->step;
    return !(a.st->fallback || step % check_frequency != 0 || step == a.st->checked_step);
}

__global__ void __launch_bounds__(256) lf_residual(StepArgs a, double* prev, double* partial, int loop,
                                                   int check_frequency, double threshold, int save) {
    if (loop && !residual_due(a, check_frequency)) return;
    residual_block(a, prev, partial, loop, check_frequency, threshold, save, (int)blockIdx.x, (int)gridDim.x);
}

// The loop's SolutionChange + SavePrev over many instances: work entry b =
// (instance, block index, the instance's block count).  Instance i reads the
// state its step wrote, args[2 i + (icur[i] ^ flip ^ 1)].
__global__ void __launch_bounds__(256) lf_residual_batch(const StepArgs* __restrict__ args,
                                                         const int* __restrict__ icur,
                                                         const BatchAux* __restrict__ aux,
                                                         const BatchBlock* __restrict__ work, int flip) {
    const BatchBlock w = work[blockIdx.x];
    StepArgs a = args[2 * w.inst + (icur[w.inst] ^ flip ^ 1)];
    const BatchAux x = aux[w.inst];
    if (x.check_frequency <= 0 || !residual_due(a, x.check_frequency)) return;
    a.step1 = 0;
    residual_block(a, x.prev, x.partial, 1, x.check_frequency, x.threshold, 1, w.first, w.count);
}

// ------------------------------------------- persistent small-instance batch

// The batched time loop of small instances in one launch: workgroup b runs up
// to K steps of instance work[b].inst on its own, its four waves taking the
// phases of each step in turn with workgroup barriers between them: the
// central update over the instance's tiles (wave w takes tiles w, w + 4, ...),
// the boundary ring (wave w takes the 64-cell groups w, w + 4, ... of the
// compact ring layout, each wave with its own LDS work area behind the
// instance's program image, staged once for the K steps), the step end
// (thread 0), and, on the steps where the loop checks convergence,
// SolutionChange + SavePrev over the single solver's partition of the
// instance into 256-thread blocks, the block sums added in block order (the
// same operations in the same order as lf_residual, so the same bits).  No
// workgroup waits for another.  Step j of instance i reads
// args[2 i + (icur[i] ^ (j & 1))]; an instance whose loop idles (done, rescan
// pending, new state outside the split division's range) stops for the rest
// of the launch, as in the graph of per-phase launches.
template <bool kBuf, class Div>
__global__ void __launch_bounds__(256) lf_batch_persist(const StepArgs* __restrict__ args,
                                                       const int* __restrict__ icur,
                                                       const BatchAux* __restrict__ aux,
                                                       const BatchBlock* __restrict__ work, int K, int img_words,
                                                       int area_slots) {
    extern __shared__ uint64_t bc_lds[];
    __shared__ int s_flag;
    const BatchBlock w = work[blockIdx.x];
    const int inst = __builtin_amdgcn_readfirstlane(w.inst);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int c0 = icur[inst];
    const BatchAux x = aux[inst];
    {
        const StepArgs a0 = args[2 * inst + c0];
        for (int i = (int)threadIdx.x; i < a0.bc_words; i += (int)blockDim.x) bc_lds[i] = a0.bc[i];
    }
    double* area = reinterpret_cast<double*>(bc_lds + img_words) + (size_t)wave * area_slots * BC_THREADS;
    for (int j = 0; j < K; ++j) {
        const StepArgs a = args[2 * inst + (c0 ^ (j & 1))];
        if (threadIdx.x == 0) s_flag = loop_idle(a) ? 0 : 1;
        __syncthreads();
        if (!s_flag) break;
        for (int t = wave; t < w.count; t += 4) stencil_tile<true, kBuf, Div>(a, 0.0, 1, w.H, w.nstrips, t);
        __syncthreads();
        {
            const double dt = step_dt(a, 0.0, 1);
            const int nring = ring_lanes_compact(a.L, a.Mn);
            for (int g = wave; g * BC_THREADS < nring; g += 4) {
                const RingLane rl = ring_lane(a, g * BC_THREADS + (int)(threadIdx.x & (BC_THREADS - 1)), false);
                bc_ring_block(a, dt, 1, bc_lds, area, rl);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            finish_instance(a, read_loop(a.st));
            s_flag = (x.check_frequency > 0 && residual_due(a, x.check_frequency)) ? 1 : 0;
        }
        __syncthreads();
        if (s_flag) {
            // The state this step wrote is the u_in of the other orientation.
            StepArgs an = args[2 * inst + (c0 ^ ((j + 1) & 1))];
            an.step1 = 0;
            const int nblk = residual_blocks(an.L, an.Mn);
            double gd = 0.0, gc = 0.0;
            for (int b = 0; b < nblk; ++b) {
                double bd, bc;
                residual_sums(an, x.prev, 1, b, nblk, bd, bc);
                gd += bd;
                gc += bc;
            }
            if (threadIdx.x == 0) residual_result(an, gd, gc, 1, x.threshold);
        }
        __syncthreads();
    }
}

// Max wave speed of the state u_in over the interior (or of the IC physical
// arrays on the first step).
__global__ void __launch_bounds__(256) lf_max_speed(StepArgs a, unsigned long long* out) {
    const int64_t n = (int64_t)a.L * a.Mn;
    uint64_t key = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int l = 1 + (int)(c / a.Mn), m = 1 + (int)(c % a.Mn);
        double ph[10];
        start_phys(a, l, m, ph);
        const uint64_t k = speed_key(wave_speed_ph(ph, a.gamma));
        key = (k > key) ? k : key;
    }
    key = wave_max_u64(key);
    if ((threadIdx.x & 63) == 0 && key != 0) atomicMax(out, (unsigned long long)key);
}

// Physical fields of the state (10 fields, ghosts 0): the arrays the reference
// keeps in Fields after a step.
__global__ void __launch_bounds__(256) lf_phys_fields(StepArgs a, double* out) {
    const int cols = a.Mn + 2;
    const int64_t n = (int64_t)(a.L + 2) * cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int l = (int)(i / cols), m = (int)(i % cols);
        double ph[10];
        if (l >= 1 && l <= a.L && m >= 1 && m <= a.Mn) {
            start_phys(a, l, m, ph);
        } else {
            for (int k = 0; k < 10; ++k) ph[k] = 0.0;
        }
        const int64_t j = at(a, l, m);
        for (int k = 0; k < 10; ++k) out[(int64_t)k * a.fstride + j] = ph[k];
    }
}

// The same physical fields for the rows r0 .. r0 + nrows - 1 only (ghost
// rows and columns 0), into out: field k's row r0 + i at
// out[(k * nrows + i) * pitch], in the strip layout of the state.
__global__ void __launch_bounds__(256) lf_phys_rows(StepArgs a, int r0, int nrows, double* out) {
    const int cols = a.Mn + 2;
    const int64_t n = (int64_t)nrows * cols;
    const int64_t fs = (int64_t)nrows * a.pitch;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), m = (int)(i % cols);
        const int l = r0 + r;
        double ph[10];
        if (l >= 1 && l <= a.L && m >= 1 && m <= a.Mn) {
            start_phys(a, l, m, ph);
        } else {
            for (int k = 0; k < 10; ++k) ph[k] = 0.0;
        }
        const int64_t j = (int64_t)r * a.pitch + phys_col(m);
        for (int k = 0; k < 10; ++k) out[(int64_t)k * fs + j] = ph[k];
    }
}

// Verification::ComputeIntegrals (src/verification.cpp:16-45): per-row sums
// in m order, rows combined in l order by the last block.  rowsum: 7*(L+2).
__global__ void __launch_bounds__(64) lf_integrals(StepArgs a, double* rowsum, double* out) {
    const int l = 1 + (int)blockIdx.x;
    __shared__ double sh[7][64];
    const double area = a.dr[l] * a.dz;
    const int ks[7] = {0, 1, 2, 3, 4, 6, 5};
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int m = 1 + (int)threadIdx.x; m <= a.Mn; m += 64) {
        const int64_t i = at(a, l, m);
        for (int k = 0; k < 7; ++k) acc[k] += a.u_in[(int64_t)ks[k] * a.fstride + i] * area;
    }
    for (int k = 0; k < 7; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < 7) {
        double s = 0.0;
        for (int t = 0; t < 64; ++t) s += sh[threadIdx.x][t];
        rowsum[(int64_t)threadIdx.x * (a.L + 2) + l] = s;
    }
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned t = atomicAdd(&a.st->ticket2, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x < 7) {
        __threadfence();
        double s = 0.0;
        for (int ll = 1; ll <= a.L; ++ll)
            s += __hip_atomic_load(&rowsum[(int64_t)threadIdx.x * (a.L + 2) + ll], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out[threadIdx.x] = s;
    }
    if (last && threadIdx.x == 0) a.st->ticket2 = 0;
}

// GetMassFlux / GetThrust (src/integrals.cpp:44-112) in the reference's serial
// m order: one thread.  out[0] = mass flux, out[1] = thrust.
__global__ void lf_outlet(StepArgs a, int M_max, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int l = a.L;
    const double dr_out = a.dr[l];
    const double pi = 3.14159265358979323846;
    const double inv_8pi = 1.0 / (8.0 * pi);
    double mf = 0.0, th = 0.0;
    for (int m = 1; m <= a.Mn; ++m) {
        const int mg = m - 1;
        const double w = (mg == 0 || mg == M_max) ? 0.5 : 1.0;
        double ph[10];
        start_phys(a, l, m, ph);
        const double r = cell_r(a, l, m);
        const double f1 = ph[PH_RHO] * ph[PH_VZ];
        const double v_z2 = ph[PH_VZ] * ph[PH_VZ];
        const double H2 = ph[PH_HZ] * ph[PH_HZ] + ph[PH_HR] * ph[PH_HR] + ph[PH_HPHI] * ph[PH_HPHI];
        const double f2 = ph[PH_RHO] * v_z2 + ph[PH_P] + H2 * inv_8pi;
        mf += w * (f1 * 2.0 * pi * r) * dr_out;
        th += w * (f2 * 2.0 * pi * r) * dr_out;
    }
    out[0] = mf;
    out[1] = th;
}

// Verification::CheckRadialSymmetry (src/verification.cpp:120-146).
__global__ void __launch_bounds__(256) lf_radial_sym(StepArgs a, unsigned long long* out) {
    const int64_t n = (int64_t)a.L * (a.Mn - 1);
    uint64_t key = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int l = 1 + (int)(c / (a.Mn - 1)), m = 1 + (int)(c % (a.Mn - 1));
        double pa[10], pb[10];
        start_phys(a, l, m, pa);
        start_phys(a, l, m + 1, pb);
        const double x = pa[PH_RHO], y = pb[PH_RHO];
        const double mean = 0.5 * (x + y);
        if (mean > 1.0e-15) {
            const uint64_t k = speed_key(fabs(y - x) / mean);
            key = (k > key) ? k : key;
        }
    }
    key = wave_max_u64(key);
    if ((threadIdx.x & 63) == 0 && key != 0) atomicMax(out, (unsigned long long)key);
}

// y[l] = div_recip(2 d[l]): the refined reciprocals of the per-row radial
// divisors 2 dr_l (src/solver.cpp:97), made once per solver on the device
// because they come from the device's v_rcp_f64.
__global__ void lf_recip_table(const double* d, double* y, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) y[i] = div_recip(2 * d[i]);
}

// Physical values of one cell, 0 on the ghost ring: what the driver's
// checkpoint line reads from Fields (src/main.cpp:259-264).
__global__ void lf_probe(StepArgs a, int l, int m, double* out) {
    if (threadIdx.x != 0) return;
    double ph[10];
    if (l >= 1 && l <= a.L && m >= 1 && m <= a.Mn) {
        start_phys(a, l, m, ph);
    } else {
        for (int k = 0; k < 10; ++k) ph[k] = 0.0;
    }
    for (int k = 0; k < 10; ++k) out[k] = ph[k];
}

// Fill kernel (used instead of memset nodes inside captured graphs).
__global__ void lf_fill(double* p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------- launchers

namespace {
// Dynamic LDS of a boundary workgroup: the program image and the work area.
size_t bc_lds_bytes(int img_words, int area_slots) {
    return ((size_t)img_words + (size_t)area_slots * BC_THREADS) * 8;
}

template <class Div>
hipError_t launch_simple(const StepArgs& a, double dt, int loop, int speed, hipStream_t s) {
    const dim3 blk(64, 4);
    const dim3 grd((unsigned)((a.Mn - 2 + 63) / 64), (unsigned)((a.L + 3) / 4));
    if (a.step1) {
        if (speed) hipLaunchKernelGGL((lf_stencil_simple<true, true, Div>), grd, blk, 0, s, a, dt, loop);
        else hipLaunchKernelGGL((lf_stencil_simple<true, false, Div>), grd, blk, 0, s, a, dt, loop);
    } else {
        if (speed) hipLaunchKernelGGL((lf_stencil_simple<false, true, Div>), grd, blk, 0, s, a, dt, loop);
        else hipLaunchKernelGGL((lf_stencil_simple<false, false, Div>), grd, blk, 0, s, a, dt, loop);
    }
    return hipGetLastError();
}

template <class Div>
hipError_t launch_tiled(const StepArgs& a, int H, int order, bool buffers, double dt, int loop, int speed,
                        hipStream_t s) {
    // Strips holding central columns 2 .. M_max (strip s: columns 62 s .. 62 s + 61).
    const int nstrips = (a.Mn - 1) / kStripCols + 1;
    const int nchunks = (a.L + H - 1) / H;
    const int ntiles = nstrips * nchunks;
    const int ngroups = (nchunks + 3) / 4;  // order 2: chunk groups of four
    const int max_cnt = (nstrips + 7) / 8;  // order 3: strips of the XCD with the most
    const unsigned blocks = order == 0   ? (unsigned)((ntiles + 3) / 4)
                            : order == 1 ? (unsigned)(8 * ((nchunks + 7) / 8) * ((nstrips + 3) / 4))
                            : order == 3 ? (unsigned)(8 * ((max_cnt * nchunks + 3) / 4))
                                         : (unsigned)(8 * ((ngroups + 7) / 8) * nstrips);
    // Buffer addressing needs every row/field offset of a state below 2^31 bytes.
    const bool buf = buffers && (double)a.fstride * 8.0 * 8.0 < 2147483648.0;
    if (buf) {
        if (speed) hipLaunchKernelGGL((lf_stencil_tiled<true, true, Div>), dim3(blocks), dim3(256), 0, s, a, dt, loop, H, nstrips, ntiles, order);
        else hipLaunchKernelGGL((lf_stencil_tiled<false, true, Div>), dim3(blocks), dim3(256), 0, s, a, dt, loop, H, nstrips, ntiles, order);
    } else {
        if (speed) hipLaunchKernelGGL((lf_stencil_tiled<true, false, Div>), dim3(blocks), dim3(256), 0, s, a, dt, loop, H, nstrips, ntiles, order);
        else hipLaunchKernelGGL((lf_stencil_tiled<false, false, Div>), dim3(blocks), dim3(256), 0, s, a, dt, loop, H, nstrips, ntiles, order);
    }
    return hipGetLastError();
}

int auto_rows(const StepArgs&) {
    // Short tiles: at 6400x2560 (42 strips) H = 8 measured fastest, ahead of
    // 24, 67 and 134 (a single round of waves) although each tile reads two
    // extra rows: the GPU then works on a narrow band of rows at a time.
    return 8;
}
}  // namespace

hipError_t launch_stencil(const StepArgs& a, const StencilCfg& c, double dt, int loop, int speed, hipStream_t s) {
    if (a.step1 || c.kind == STENCIL_SIMPLE) {
        // The first step reads the IC physical arrays: always the simple kernel.
        if (c.fast_div && !a.step1) return launch_simple<DivConst>(a, dt, loop, speed, s);
        return launch_simple<DivExact>(a, dt, loop, speed, s);
    }
    const int H = c.tile_rows > 0 ? c.tile_rows : auto_rows(a);
    if (c.fast_div) return launch_tiled<DivConst>(a, H, c.order, c.buffers, dt, loop, speed, s);
    return launch_tiled<DivExact>(a, H, c.order, c.buffers, dt, loop, speed, s);
}

hipError_t launch_bc(const StepArgs& a, double dt, int loop, int speed, hipStream_t s) {
    const int nring = ring_lanes(a.L, a.Mn);
    hipLaunchKernelGGL(lf_bc_ring, dim3((unsigned)((nring + BC_THREADS - 1) / BC_THREADS)), dim3(BC_THREADS),
                       bc_lds_bytes(a.bc_words, a.bc_area), s, a, dt, loop, speed);
    if (loop) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(lf_finish_one, dim3(1), dim3(1), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_stencil_batch(const StepArgs* args, const int* icur, const BatchBlock* work, int nwork, int grid,
                                int flip, bool fast_div, bool buf, double dt, int loop, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    const dim3 g((unsigned)std::max(1, std::min(grid > 0 ? grid : nwork, nwork)));
    if (buf) {
        if (fast_div) hipLaunchKernelGGL((lf_stencil_tiled_batch<true, true, DivConst>), g, dim3(256), 0, s, args, icur, work, nwork, flip, dt, loop);
        else hipLaunchKernelGGL((lf_stencil_tiled_batch<true, true, DivExact>), g, dim3(256), 0, s, args, icur, work, nwork, flip, dt, loop);
    } else {
        if (fast_div) hipLaunchKernelGGL((lf_stencil_tiled_batch<true, false, DivConst>), g, dim3(256), 0, s, args, icur, work, nwork, flip, dt, loop);
        else hipLaunchKernelGGL((lf_stencil_tiled_batch<true, false, DivExact>), g, dim3(256), 0, s, args, icur, work, nwork, flip, dt, loop);
    }
    return hipGetLastError();
}

hipError_t launch_bc_batch(const StepArgs* args, const int* icur, const BatchBlock* work, int nwork, int flip,
                           int img_words, int area_slots, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    const size_t lds = bc_lds_bytes(img_words, area_slots);
    hipLaunchKernelGGL(lf_bc_ring_batch, dim3((unsigned)nwork), dim3(BC_THREADS), lds, s, args, icur, work, flip, 1,
                       img_words);
    return hipGetLastError();
}

hipError_t launch_finish_batch(const StepArgs* args, const int* icur, int n, int flip, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lf_finish_batch, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, args, icur, n, flip);
    return hipGetLastError();
}

hipError_t launch_residual_batch(const StepArgs* args, const int* icur, const BatchAux* aux, const BatchBlock* work,
                                 int nwork, int flip, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(lf_residual_batch, dim3((unsigned)nwork), dim3(256), 0, s, args, icur, aux, work, flip);
    return hipGetLastError();
}

size_t persist_lds_bytes(int img_words, int area_slots) {
    return ((size_t)img_words + (size_t)4 * area_slots * BC_THREADS) * 8;
}

hipError_t launch_batch_persist(const StepArgs* args, const int* icur, const BatchAux* aux, const BatchBlock* work,
                                int nwork, int K, bool fast_div, bool buf, int img_words, int area_slots,
                                hipStream_t s) {
    if (nwork <= 0 || K <= 0) return hipSuccess;
    const size_t lds = persist_lds_bytes(img_words, area_slots);
    const dim3 g((unsigned)nwork), b(256);
    // Dynamic LDS above 64 KiB has to be allowed per kernel.
    auto launch = [&](auto kernel) -> hipError_t {
        if (lds > 65536) {
            const hipError_t e =
                hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kernel, g, b, lds, s, args, icur, aux, work, K, img_words, area_slots);
        return hipGetLastError();
    };
    if (buf) return fast_div ? launch(lf_batch_persist<true, DivConst>) : launch(lf_batch_persist<true, DivExact>);
    return fast_div ? launch(lf_batch_persist<false, DivConst>) : launch(lf_batch_persist<false, DivExact>);
}

hipError_t launch_speed_fallback(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(lf_speed_fallback, dim3(256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_residual(const StepArgs& a, double* prev, double* partial, int loop, int check_frequency,
                           double threshold, int save, hipStream_t s) {
    hipLaunchKernelGGL(lf_residual, dim3((unsigned)residual_blocks(a.L, a.Mn)), dim3(256), 0, s, a, prev, partial, loop,
                       check_frequency, threshold, save);
    return hipGetLastError();
}

hipError_t launch_max_speed(const StepArgs& a, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_max_speed, dim3(1024), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

hipError_t launch_phys_fields(const StepArgs& a, double* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_phys_fields, dim3(1024), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

hipError_t launch_phys_rows(const StepArgs& a, int r0, int nrows, double* out, hipStream_t s) {
    const long long n = (long long)nrows * (a.Mn + 2);
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(lf_phys_rows, dim3((unsigned)(blocks < 1024 ? blocks : 1024)), dim3(256), 0, s, a, r0, nrows,
                       out);
    return hipGetLastError();
}

hipError_t launch_integrals(const StepArgs& a, double* rowsum, double* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_integrals, dim3((unsigned)a.L), dim3(64), 0, s, a, rowsum, out);
    return hipGetLastError();
}

hipError_t launch_outlet(const StepArgs& a, int M_max, double* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_outlet, dim3(1), dim3(64), 0, s, a, M_max, out);
    return hipGetLastError();
}

hipError_t launch_radial_sym(const StepArgs& a, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_radial_sym, dim3(1024), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

hipError_t launch_probe(const StepArgs& a, int l, int m, double* out, hipStream_t s) {
    hipLaunchKernelGGL(lf_probe, dim3(1), dim3(64), 0, s, a, l, m, out);
    return hipGetLastError();
}

hipError_t launch_recip_table(const double* d, double* y, int n, hipStream_t s) {
    hipLaunchKernelGGL(lf_recip_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, y, n);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, long long n, double v, hipStream_t s) {
    hipLaunchKernelGGL(lf_fill, dim3(1024), dim3(256), 0, s, p, (int64_t)n, v);
    return hipGetLastError();
}

}  // namespace mpd
