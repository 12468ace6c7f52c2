// lf_device.h -- device-side math of the Lax-Friedrichs step shared by the
// gfx950 kernels.  Every function restates a reference formula in the
// reference's exact operation order (IEEE binary64, no contraction: the
// library is compiled with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lf_types.h"

namespace mpd {

// Kernel parameters of one step (all device pointers).
struct StepArgs {
    int L;          // L_max: interior rows 1..L
    int Mn;         // M_max + 1: interior columns 1..Mn
    int pitch;      // row pitch in doubles
    int64_t fstride;  // distance between two fields of one state buffer, in doubles
    const double* u_in;   // u0 (start of step), 8 fields
    double* u_out;        // u (end of step), 8 fields
    double* rowbuf;       // central u of rows 1 and L: [2][8][pitch]
    const double* phys_in;  // step-1 physical fields (10, same layout) or nullptr
    const double* rin;    // per row l = 0..L+1: r_inner(z_l)
    const double* rout;   // r_outer(z_l)
    const double* drin;   // d r_inner / dz
    const double* drout;  // d r_outer / dz
    const double* dr;     // dr[l] = (r_outer - r_inner) / M_max
    const double* dr2inv; // div_recip(2 dr[l]), computed on the device (lf_recip_table)
    double gamma, gm1, beta, H_z0, r_0, dz, dy, dz2;
    double cfl_dx, growth, dt_min, dt_max, T;
    int adaptive;
    int step1;            // start-of-step physical comes from phys_in
    int bc_words;         // size of the BC program's LDS image in 8-byte words (bc_image_words)
    int bc_area;          // work-area slots per lane of the boundary kernel (bc_area_slots)
    MpdLoopState* st;
    const uint64_t* bc;   // the BC program's image (BcImage layout, bc_words words)
    const double* ringgeo;  // per boundary cell geometry (RingGeo layout, ring_compact_index order)
};

// One workgroup's entry in a batched launch's work list: its instance and,
// for the stencil, the first of its four tiles, the instance's tile count,
// tile rows and strip count; for the boundary and residual kernels, the
// block index within the instance and the instance's block count.
struct BatchBlock {
    int inst;
    int first;
    int count;
    int H;
    int nstrips;
};

// Per-instance convergence data of a batched loop (check_frequency 0: no
// convergence check for that instance).
struct BatchAux {
    double* prev;
    double* partial;
    double threshold;
    int check_frequency;
    int pad;
};

// Image of a BC program, built on the host and staged into LDS by the
// boundary kernels: the counts, then the segments in use (faces in the order
// L_LO, L_HI, M_LO, M_HI, each face's in config order), then the expression
// code in use.  Its size follows the program, so the boundary kernels load
// exactly bc_image_words words in one round trip and take the image, with
// the work area behind it, in dynamic LDS.
struct BcImage {
    int32_t nseg[4];
    int32_t ncode;
    int32_t first[4];  // index of face f's first segment
    int32_t nsegs;
    __host__ __device__ __forceinline__ const MpdSeg& seg(int f, int s) const {
        return reinterpret_cast<const MpdSeg*>(this + 1)[first[f] + s];
    }
    __host__ __device__ __forceinline__ const MpdInstr& code(int i) const {
        return reinterpret_cast<const MpdInstr*>(reinterpret_cast<const MpdSeg*>(this + 1) + nsegs)[i];
    }
};
constexpr int kBcImageHeadWords = 5;
static_assert(sizeof(BcImage) == 8 * kBcImageHeadWords, "image header layout");
static_assert(sizeof(MpdSeg) % 8 == 0 && sizeof(MpdInstr) == 16, "image word layout");

__host__ __device__ inline int bc_image_words(const MpdBcProgram& p) {
    return kBcImageHeadWords + (int)(sizeof(MpdSeg) / 8) * (p.nseg[0] + p.nseg[1] + p.nseg[2] + p.nseg[3]) +
           2 * p.ncode;
}

// Writes the image of p into out (bc_image_words(p) words).
inline void build_bc_image(const MpdBcProgram& p, uint64_t* out) {
    BcImage h{};
    int first = 0;
    for (int f = 0; f < 4; ++f) {
        h.nseg[f] = p.nseg[f];
        h.first[f] = first;
        first += p.nseg[f];
    }
    h.ncode = p.ncode;
    h.nsegs = first;
    char* o = reinterpret_cast<char*>(out);
    __builtin_memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    for (int f = 0; f < 4; ++f) {
        __builtin_memcpy(o, &p.seg[f][0], sizeof(MpdSeg) * p.nseg[f]);
        o += sizeof(MpdSeg) * p.nseg[f];
    }
    __builtin_memcpy(o, &p.code[0], sizeof(MpdInstr) * p.ncode);
}

// Per-lane work area of the boundary kernel, in slots of one double: the
// expression symbol table, then the evaluation stack, as deep as the
// program's deepest expression needs.
enum { WA_SYM = 0, WA_STK = WA_SYM + MPD_BS_COUNT };
// Expressions at most this deep evaluate on a stack held in registers.
constexpr int kRegStack = 4;
// Boundary-kernel lanes of the ring of an L x Mn grid.  Compact (the
// batched kernel, where many instances' workgroups must fit at once): row 1,
// row L, column 1 (l = 2..L-1), column Mn, back to back.  Grouped (a single
// grid): 64-lane groups of one kind of cell each: the interior cells of row 1 (m = 2..Mn-1), of row L,
// of column 1 (l = 2..L-1) and of column Mn, each padded to whole groups,
// then one group per corner (1,1), (1,Mn), (L,1), (L,Mn) with its first lane
// only.  A 64-lane wave then evaluates one face's conditions, and a corner
// (whose evaluation may also run an earlier face's chain at its neighbour)
// runs alone: lanes of different kinds in one wave would take their paths
// one after another.
__host__ __device__ inline int ring_lanes_compact(int L, int Mn) { return 2 * Mn + 2 * (L - 2); }

__host__ __device__ inline bool ring_cell_compact(int idx, int L, int Mn, int& l, int& m) {
    if (idx < Mn) { l = 1; m = idx + 1; }
    else if (idx < 2 * Mn) { l = L; m = idx - Mn + 1; }
    else if (idx < 2 * Mn + L - 2) { l = idx - 2 * Mn + 2; m = 1; }
    else { l = idx - (2 * Mn + L - 2) + 2; m = Mn; }
    return idx < ring_lanes_compact(L, Mn);
}

// Index of boundary cell (l, m) in the compact order (the inverse of
// ring_cell_compact).
__host__ __device__ inline int ring_compact_index(int l, int m, int L, int Mn) {
    if (l == 1) return m - 1;
    if (l == L) return Mn + m - 1;
    if (m == 1) return 2 * Mn + l - 2;
    return 2 * Mn + (L - 2) + l - 2;
}

// Geometry a boundary cell's face chains read, precomputed on the host per
// boundary cell (so the boundary kernel loads it with the cell states, in
// one round trip, instead of per-row wall data in dependent loads): r and
// r_z of the cell, of its row neighbour (l +- 1 inwards; the cell itself on
// a column), of its column neighbour (m +- 1 inwards) and of the diagonal
// between them, dr and z of the cell's row and of the row neighbour's.
// Stored field-major: entry e of cell c at ringgeo[e * n + c], n =
// ring_lanes_compact(L, Mn).
enum { RG_R = 0, RG_RZ, RG_DR, RG_Z, RG_R_ROW, RG_RZ_ROW, RG_DR_ROW, RG_Z_ROW, RG_R_COL, RG_RZ_COL, RG_R_DIAG,
       RG_RZ_DIAG, RG_N };

__host__ __device__ inline int ring_lanes(int L, int Mn) {
    const int g1 = ((Mn - 2) + 63) / 64 * 64, gc = ((L - 2) + 63) / 64 * 64;
    return 2 * g1 + 2 * gc + 4 * 64;
}

// Cell (l, m) of ring lane idx; false for a padding lane.
__host__ __device__ inline bool ring_cell(int idx, int L, int Mn, int& l, int& m) {
    const int n1 = Mn - 2, nc = L - 2;
    const int g1 = (n1 + 63) / 64 * 64, gc = (nc + 63) / 64 * 64;
    if (idx < g1) { l = 1; m = 2 + idx; return idx < n1; }
    idx -= g1;
    if (idx < g1) { l = L; m = 2 + idx; return idx < n1; }
    idx -= g1;
    if (idx < gc) { l = 2 + idx; m = 1; return idx < nc; }
    idx -= gc;
    if (idx < gc) { l = 2 + idx; m = Mn; return idx < nc; }
    idx -= gc;
    const int c = idx / 64;
    if (idx % 64 != 0 || c >= 4) return false;
    l = (c < 2) ? 1 : L;
    m = (c % 2 == 0) ? 1 : Mn;
    return true;
}

__host__ __device__ inline int bc_area_slots(const MpdBcProgram& p) {
    return WA_STK + (p.max_stack > 1 ? p.max_stack : 1);
}

// Physical fields in the member order of include/fields.hpp:36-41.
enum { PH_RHO = 0, PH_VZ, PH_VR, PH_VPHI, PH_E, PH_P, PH_PT, PH_HZ, PH_HR, PH_HPHI };

// Per-cell record used by the stencil: the 8 conserved values, the physical
// values its neighbours read, and r.
struct Rec {
    double u[8];
    double vz, vr, vphi, P, Hz, Hr, Hphi, r;
    double rho, p;
};

__device__ __forceinline__ double cell_r(const StepArgs& a, int l, int m) {
    // Grid::Build, src/grid.cpp:31-37: frac = m_global * dy, m_global = m - 1.
    const double frac = (double)(m - 1) * a.dy;
    return (1.0 - frac) * a.rin[l] + frac * a.rout[l];
}

__device__ __forceinline__ double cell_rz(const StepArgs& a, int l, int m) {
    const double frac = (double)(m - 1) * a.dy;
    return (1.0 - frac) * a.drin[l] + frac * a.drout[l];
}

// Fields::UpdatePhysicalFromU, src/fields.cpp:84-97.
__device__ __forceinline__ void rec_phys(Rec& c, double gm1) {
    c.rho = c.u[0] / c.r;
    c.vz = c.u[1] / c.u[0];
    c.vr = c.u[2] / c.u[0];
    c.vphi = c.u[3] / c.u[0];
    c.Hphi = c.u[5];
    c.Hz = c.u[6] / c.r;
    c.Hr = c.u[7] / c.r;
    const double e = c.u[4] / c.u[0];
    c.p = gm1 * c.rho * e;
    c.P = c.p + 0.5 * (c.Hz * c.Hz + c.Hr * c.Hr + c.Hphi * c.Hphi);
}

__device__ __forceinline__ void phys10_from_u(const double* u, double r, double gm1, double* ph) {
    ph[PH_RHO] = u[0] / r;
    ph[PH_VZ] = u[1] / u[0];
    ph[PH_VR] = u[2] / u[0];
    ph[PH_VPHI] = u[3] / u[0];
    ph[PH_HPHI] = u[5];
    ph[PH_HZ] = u[6] / r;
    ph[PH_HR] = u[7] / r;
    ph[PH_E] = u[4] / u[0];
    ph[PH_P] = gm1 * ph[PH_RHO] * ph[PH_E];
    ph[PH_PT] = ph[PH_P] + 0.5 * (ph[PH_HZ] * ph[PH_HZ] + ph[PH_HR] * ph[PH_HR] + ph[PH_HPHI] * ph[PH_HPHI]);
}

// Diagnostics::MaxWaveSpeed's per-cell term, src/diagnostics.cpp:26-36, on
// the physical values f(u) of conserved values u.
__device__ __forceinline__ double wave_speed_u(const double* u, double r, double gamma, double gm1) {
    const double rho = u[0] / r;
    const double vz = u[1] / u[0];
    const double vr = u[2] / u[0];
    const double hphi = u[5];
    const double hz = u[6] / r;
    const double hr = u[7] / r;
    const double e = u[4] / u[0];
    const double p = gm1 * rho * e;
    const double cs = sqrt(gamma * p / rho);
    const double ca = sqrt((hz * hz + hr * hr + hphi * hphi) / rho);
    const double v = sqrt(vz * vz + vr * vr);
    return v + cs + ca;
}

__device__ __forceinline__ double wave_speed_ph(const double* ph, double gamma) {
    const double rho = ph[PH_RHO];
    const double hz = ph[PH_HZ], hr = ph[PH_HR], hphi = ph[PH_HPHI];
    const double vz = ph[PH_VZ], vr = ph[PH_VR];
    const double cs = sqrt(gamma * ph[PH_P] / rho);
    const double ca = sqrt((hz * hz + hr * hr + hphi * hphi) / rho);
    const double v = sqrt(vz * vz + vr * vr);
    return v + cs + ca;
}

// Division split in two halves.  The compiler lowers the IEEE binary64
// division x / d on gfx950 (LLVM AMDGPU fdiv64) as: range scaling of x and d
// (v_div_scale_f64), y = v_rcp_f64(d) refined by two Newton steps
// (four FMAs), q = x * y, one FMA residual r = x - d q, then
// v_div_fmas_f64(r, y, q) = fma(r, y, q) and v_div_fixup_f64 for special
// operands.  div_recip is the divisor half, div_by the quotient half, so a
// divisor shared by several quotients is refined once.  The residual is
// carried negated, e = q d - x = -r (round-to-nearest is symmetric), so a
// zero x keeps its sign as v_div_fixup_f64 would give it.  For d > 0 and x
// zero or finite, and operands and quotient inside the range where the
// scaling step leaves them unchanged (|x|, |d| and |x / d| between 2^-380
// and 2^380 holds for every physical state), div_by(x, d, div_recip(d)) is
// bit-identical to x / d, which is correctly rounded.
__device__ __forceinline__ double div_recip(double d) {
    const double y0 = __builtin_amdgcn_rcp(d);
    const double e0 = __builtin_fma(-d, y0, 1.0);
    const double y1 = __builtin_fma(y0, e0, y0);
    const double e1 = __builtin_fma(-d, y1, 1.0);
    return __builtin_fma(y1, e1, y1);
}

__device__ __forceinline__ double div_by(double x, double d, double y) {
    const double q = x * y;
    const double e = __builtin_fma(q, d, -x);
    return __builtin_fma(-e, y, q);
}

// Fields::UpdatePhysicalFromU (src/fields.cpp:84-97) with the divisions by
// u_1 and by r sharing one refined reciprocal each (same bits as rec_phys
// for every state with u_1, r > 0 inside the range div_by states).
__device__ __forceinline__ void rec_phys_shared(Rec& c, double gm1) {
    const double yu = div_recip(c.u[0]);
    const double yr = div_recip(c.r);
    c.rho = div_by(c.u[0], c.r, yr);
    c.vz = div_by(c.u[1], c.u[0], yu);
    c.vr = div_by(c.u[2], c.u[0], yu);
    c.vphi = div_by(c.u[3], c.u[0], yu);
    c.Hphi = c.u[5];
    c.Hz = div_by(c.u[6], c.r, yr);
    c.Hr = div_by(c.u[7], c.r, yr);
    const double e = div_by(c.u[4], c.u[0], yu);
    c.p = gm1 * c.rho * e;
    c.P = c.p + 0.5 * (c.Hz * c.Hz + c.Hr * c.Hr + c.Hphi * c.Hphi);
}

// Whether the next step's split divisions on conserved values u are
// bit-identical to IEEE division: |u_1| in [2^-30, 2^31) and every other
// value zero or |u_k| in [2^-100, 2^61) (binary exponents from
// v_frexp_exp_i32_f64, which gives 0 for a zero).  With every radius in
// [2^-8, 2^8] (checked on the host) the physical values, their flux
// products and every difference the step divides lie between 2^-331 and
// 2^221 (or are zero), inside the range div_by states.  A step whose new
// state fails the test flags the loop state; the host continues with IEEE
// division from the next step on.
__device__ __forceinline__ bool split_div_safe(const double* u) {
    bool ok = (unsigned)(__builtin_amdgcn_frexp_exp(u[0]) + 29) <= 60u;
#pragma unroll
    for (int k = 1; k < 8; ++k) ok = ok && (unsigned)(__builtin_amdgcn_frexp_exp(u[k]) + 99) <= 160u;
    return ok;
}

// Flags the loop state when any lane of the wave saw a value outside the
// range (one atomic per wave, and only then).
__device__ __forceinline__ void flag_range(MpdLoopState* st, bool bad) {
    if (__builtin_amdgcn_ballot_w64(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(&st->range_bits, 1u);
}

struct DivExact {
    __device__ __forceinline__ double operator()(double x, double c, double) const { return x / c; }
};
// Division by the per-row constants 2 dz and 2 dr_l given ci = div_recip(c).
struct DivConst {
    __device__ __forceinline__ double operator()(double x, double c, double ci) const {
        return div_by(x, c, ci);
    }
};

// Solver::ComputeCentralUpdate's cell update, src/solver.cpp:94-150.
// N/S: l+1 / l-1; E/W: m+1 / m-1; C: the cell; dz2 = 2*dz, dr2 = 2*dr_l.
template <class Div>
__device__ __forceinline__ void lf_cell(const Rec& N, const Rec& S, const Rec& E, const Rec& W,
                                        const Rec& C, double dt, double dz2, double dz2i,
                                        double dr2, double dr2i, double* o) {
    const Div D;
    o[0] = 0.25 * (N.u[0] + S.u[0] + E.u[0] + W.u[0]) +
           dt * (D(-(N.u[0] * N.vz - S.u[0] * S.vz), dz2, dz2i) -
                 D(E.u[0] * E.vr - W.u[0] * W.vr, dr2, dr2i));

    o[1] = 0.25 * (N.u[1] + S.u[1] + E.u[1] + W.u[1]) +
           dt * (D((N.Hz * N.Hz - N.P) * N.r - (S.Hz * S.Hz - S.P) * S.r, dz2, dz2i) +
                 D(E.Hz * E.Hr * E.r - W.Hz * W.Hr * W.r, dr2, dr2i) -
                 D(N.u[1] * N.vz - S.u[1] * S.vz, dz2, dz2i) -
                 D(E.u[1] * E.vr - W.u[1] * W.vr, dr2, dr2i));

    o[2] = 0.25 * (N.u[2] + S.u[2] + E.u[2] + W.u[2]) +
           dt * ((C.rho * C.vphi * C.vphi + C.P - C.Hphi * C.Hphi) +
                 D(N.Hz * N.Hr * N.r - S.Hz * S.Hr * S.r, dz2, dz2i) +
                 D((E.Hr * E.Hr - E.P) * E.r - (W.Hr * W.Hr - W.P) * W.r, dr2, dr2i) -
                 D(N.u[2] * N.vz - S.u[2] * S.vz, dz2, dz2i) -
                 D(E.u[2] * E.vr - W.u[2] * W.vr, dr2, dr2i));

    o[3] = 0.25 * (N.u[3] + S.u[3] + E.u[3] + W.u[3]) +
           dt * ((-C.rho * C.vr * C.vphi + C.Hphi * C.Hr) +
                 D(N.Hphi * N.Hz * N.r - S.Hphi * S.Hz * S.r, dz2, dz2i) +
                 D(E.Hphi * E.Hr * E.r - W.Hphi * W.Hr * W.r, dr2, dr2i) -
                 D(N.u[3] * N.vz - S.u[3] * S.vz, dz2, dz2i) -
                 D(E.u[3] * E.vr - W.u[3] * W.vr, dr2, dr2i));

    o[4] = 0.25 * (N.u[4] + S.u[4] + E.u[4] + W.u[4]) +
           dt * (-C.p * (D(N.vz * N.r - S.vz * S.r, dz2, dz2i) +
                         D(E.vr * E.r - W.vr * W.r, dr2, dr2i)) -
                 D(N.u[4] * N.vz - S.u[4] * S.vz, dz2, dz2i) -
                 D(E.u[4] * E.vr - W.u[4] * W.vr, dr2, dr2i));

    o[5] = 0.25 * (N.u[5] + S.u[5] + E.u[5] + W.u[5]) +
           dt * (D(N.Hz * N.vphi - S.Hz * S.vphi, dz2, dz2i) +
                 D(E.Hr * E.vphi - W.Hr * W.vphi, dr2, dr2i) -
                 D(N.u[5] * N.vz - S.u[5] * S.vz, dz2, dz2i) -
                 D(E.u[5] * E.vr - W.u[5] * W.vr, dr2, dr2i));

    o[6] = 0.25 * (N.u[6] + S.u[6] + E.u[6] + W.u[6]) +
           dt * (D(E.Hr * E.vz * E.r - W.Hr * W.vz * W.r, dr2, dr2i) -
                 D(E.u[6] * E.vr - W.u[6] * W.vr, dr2, dr2i));

    o[7] = 0.25 * (N.u[7] + S.u[7] + E.u[7] + W.u[7]) +
           dt * (D(N.Hz * N.vr * N.r - S.Hz * S.vr * S.r, dz2, dz2i) -
                 D(N.u[7] * N.vz - S.u[7] * S.vz, dz2, dz2i));
}

// Radial record of a cell: what lf_cell reads from an m +- 1 neighbour,
// the radial flux products (each in the reference's operation order) and
// then the 8 conserved values (entry XB_U + k).
enum {
    XB_U0VR = 0, XB_HZHRR, XB_U1VR, XB_HRHRPR, XB_U2VR, XB_HPHIHRR, XB_U3VR, XB_VRR,
    XB_U4VR, XB_HRVPHI, XB_U5VR, XB_HRVZR, XB_U6VR, XB_U, XB_N = XB_U + 8
};

// Solver::ComputeCentralUpdate's cell update (src/solver.cpp:94-150) with the
// m +- 1 neighbours given as radial records: nb.E(i) / nb.W(i) is entry i of
// the E / W neighbour's radial record.  Same operations in the same order as
// lf_cell.  The scheduling barriers between the eight fields keep the
// neighbour values of one field from being fetched while another is computed
// (bounded register live ranges: the tiled stencil then needs no spills).
template <class Div, class NB>
__device__ __forceinline__ void lf_cell_rad(const Rec& N, const Rec& S, const Rec& C, const NB& nb,
                                            double dt, double dz2, double dz2i, double dr2, double dr2i,
                                            double* o) {
    const Div D;
#define MPD_FIELD_FENCE __builtin_amdgcn_sched_barrier(0)
    o[0] = 0.25 * (N.u[0] + S.u[0] + nb.E(XB_U + 0) + nb.W(XB_U + 0)) +
           dt * (D(-(N.u[0] * N.vz - S.u[0] * S.vz), dz2, dz2i) -
                 D(nb.E(XB_U0VR) - nb.W(XB_U0VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[1] = 0.25 * (N.u[1] + S.u[1] + nb.E(XB_U + 1) + nb.W(XB_U + 1)) +
           dt * (D((N.Hz * N.Hz - N.P) * N.r - (S.Hz * S.Hz - S.P) * S.r, dz2, dz2i) +
                 D(nb.E(XB_HZHRR) - nb.W(XB_HZHRR), dr2, dr2i) -
                 D(N.u[1] * N.vz - S.u[1] * S.vz, dz2, dz2i) -
                 D(nb.E(XB_U1VR) - nb.W(XB_U1VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[2] = 0.25 * (N.u[2] + S.u[2] + nb.E(XB_U + 2) + nb.W(XB_U + 2)) +
           dt * ((C.rho * C.vphi * C.vphi + C.P - C.Hphi * C.Hphi) +
                 D(N.Hz * N.Hr * N.r - S.Hz * S.Hr * S.r, dz2, dz2i) +
                 D(nb.E(XB_HRHRPR) - nb.W(XB_HRHRPR), dr2, dr2i) -
                 D(N.u[2] * N.vz - S.u[2] * S.vz, dz2, dz2i) -
                 D(nb.E(XB_U2VR) - nb.W(XB_U2VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[3] = 0.25 * (N.u[3] + S.u[3] + nb.E(XB_U + 3) + nb.W(XB_U + 3)) +
           dt * ((-C.rho * C.vr * C.vphi + C.Hphi * C.Hr) +
                 D(N.Hphi * N.Hz * N.r - S.Hphi * S.Hz * S.r, dz2, dz2i) +
                 D(nb.E(XB_HPHIHRR) - nb.W(XB_HPHIHRR), dr2, dr2i) -
                 D(N.u[3] * N.vz - S.u[3] * S.vz, dz2, dz2i) -
                 D(nb.E(XB_U3VR) - nb.W(XB_U3VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[4] = 0.25 * (N.u[4] + S.u[4] + nb.E(XB_U + 4) + nb.W(XB_U + 4)) +
           dt * (-C.p * (D(N.vz * N.r - S.vz * S.r, dz2, dz2i) +
                         D(nb.E(XB_VRR) - nb.W(XB_VRR), dr2, dr2i)) -
                 D(N.u[4] * N.vz - S.u[4] * S.vz, dz2, dz2i) -
                 D(nb.E(XB_U4VR) - nb.W(XB_U4VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[5] = 0.25 * (N.u[5] + S.u[5] + nb.E(XB_U + 5) + nb.W(XB_U + 5)) +
           dt * (D(N.Hz * N.vphi - S.Hz * S.vphi, dz2, dz2i) +
                 D(nb.E(XB_HRVPHI) - nb.W(XB_HRVPHI), dr2, dr2i) -
                 D(N.u[5] * N.vz - S.u[5] * S.vz, dz2, dz2i) -
                 D(nb.E(XB_U5VR) - nb.W(XB_U5VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[6] = 0.25 * (N.u[6] + S.u[6] + nb.E(XB_U + 6) + nb.W(XB_U + 6)) +
           dt * (D(nb.E(XB_HRVZR) - nb.W(XB_HRVZR), dr2, dr2i) -
                 D(nb.E(XB_U6VR) - nb.W(XB_U6VR), dr2, dr2i));
    MPD_FIELD_FENCE;
    o[7] = 0.25 * (N.u[7] + S.u[7] + nb.E(XB_U + 7) + nb.W(XB_U + 7)) +
           dt * (D(N.Hz * N.vr * N.r - S.Hz * S.vr * S.r, dz2, dz2i) -
                 D(N.u[7] * N.vz - S.u[7] * S.vz, dz2, dz2i));
#undef MPD_FIELD_FENCE
}

// Entry i of the radial record of cell c.
__device__ __forceinline__ double radial_entry(const Rec& c, int i) {
    switch (i) {
        case XB_U0VR: return c.u[0] * c.vr;
        case XB_HZHRR: return c.Hz * c.Hr * c.r;
        case XB_U1VR: return c.u[1] * c.vr;
        case XB_HRHRPR: return (c.Hr * c.Hr - c.P) * c.r;
        case XB_U2VR: return c.u[2] * c.vr;
        case XB_HPHIHRR: return c.Hphi * c.Hr * c.r;
        case XB_U3VR: return c.u[3] * c.vr;
        case XB_VRR: return c.vr * c.r;
        case XB_U4VR: return c.u[4] * c.vr;
        case XB_HRVPHI: return c.Hr * c.vphi;
        case XB_U5VR: return c.u[5] * c.vr;
        case XB_HRVZR: return c.Hr * c.vz * c.r;
        case XB_U6VR: return c.u[6] * c.vr;
        default: return c.u[i - XB_U];
    }
}

// Max of non-negative doubles through their bit patterns; NaN is skipped as
// std::max(local_max, v) skips it in src/diagnostics.cpp:36.
__device__ __forceinline__ uint64_t speed_key(double s) {
    return (s == s) ? (uint64_t)__double_as_longlong(s) : 0ull;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

}  // namespace mpd
