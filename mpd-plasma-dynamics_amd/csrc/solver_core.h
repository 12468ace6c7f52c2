// solver_core.h -- the device-resident solver behind the C ABI of
// include/mpd_lf.h (the reference's Solver, include/solver.hpp:24-53, and the
// reductions its driver runs between steps), shared by the single-instance
// entry points (solver.hip) and the batched sweep (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpd_lf.h"
#include "../host/config.h"
#include "../host/expr.h"
#include "../host/setup.h"
#include "../host/vtk.h"
#include "lf_launch.h"


using mpd::StepArgs;

// Message of the calling thread's last failure (mpd_last_error).
extern thread_local std::string g_last_error;

namespace mpd_detail {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline const char* const kBcSymNames[MPD_BS_COUNT] = {
    "gamma", "beta", "H_z0", "r_0", "dz", "dr", "r", "r_z", "z", "r_nb", "r_z_nb",
    "rho_nb", "v_z_nb", "v_r_nb", "v_phi_nb", "e_nb", "H_z_nb", "H_r_nb", "H_phi_nb",
    "rho", "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"};
inline const char* const kCondNames[8] = {"rho", "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"};

// PerFieldBC's constructor checks (src/bcs/per_field_bc.cpp:188-243) and the
// segment range resolution of FaceBC::Apply (src/face_bc.cpp:104-129), for one
// rank owning every face; expressions compiled to bytecode.
inline void CompileBc(const mpd::SimConfig& cfg, MpdBcProgram& bc) {
    std::memset(&bc, 0, sizeof(bc));
    const mpd::BCFaceConfig* faces[4] = {&cfg.bc_l_lo, &cfg.bc_l_hi, &cfg.bc_m_lo, &cfg.bc_m_hi};
    std::vector<std::string> names(kBcSymNames, kBcSymNames + MPD_BS_COUNT);
    int ncode = 0;
    for (int f = 0; f < 4; ++f) {
        const auto& segs = faces[f]->segments;
        if ((int)segs.size() > MPD_MAX_SEGS_PER_FACE)
            throw std::runtime_error("boundary_conditions: at most " + std::to_string(MPD_MAX_SEGS_PER_FACE) +
                                     " segments per face are supported");
        bc.nseg[f] = (int)segs.size();
        const bool is_l = (f == MPD_L_LO || f == MPD_L_HI);
        const int face_hi = is_l ? cfg.M_max : cfg.L_max - 1;
        for (size_t si = 0; si < segs.size(); ++si) {
            const mpd::BCSegmentConfig& sc = segs[si];
            MpdSeg& g = bc.seg[f][si];
            const int g_lo = sc.global_lo < 0 ? 0 : sc.global_lo;
            const int g_hi = sc.global_hi < 0 ? face_hi : sc.global_hi;
            const int c_lo = std::max(g_lo, 0), c_hi = std::min(g_hi, face_hi);
            int lo = c_lo + 1, hi = c_hi + 1;
            if (c_lo > c_hi) {
                lo = 1;
                hi = 0;
            } else if (f == MPD_M_LO) {
                lo = std::max(lo, 2);
                hi = std::min(hi, cfg.L_max - 1);
            }
            g.lo = lo;
            g.hi = hi;
            const mpd::FieldCond* fc[8] = {&sc.rho, &sc.v_z, &sc.v_r, &sc.v_phi, &sc.H_z, &sc.H_r, &sc.H_phi, &sc.e};
            const bool axis_main = sc.rho.type == mpd::FieldCondType::AxisLF ||
                                   sc.v_z.type == mpd::FieldCondType::AxisLF ||
                                   sc.e.type == mpd::FieldCondType::AxisLF ||
                                   sc.H_z.type == mpd::FieldCondType::AxisLF;
            if (axis_main && f != MPD_M_LO)
                throw std::runtime_error(
                    "PerFieldBC: AxisLF condition is only valid on the M_LO face (inner boundary / axis of symmetry)");
            const int no_axis[4] = {2, 3, 5, 6};  // v_r, v_phi, H_r, H_phi
            for (int k : no_axis) {
                if (fc[k]->type == mpd::FieldCondType::AxisLF)
                    throw std::runtime_error(std::string("PerFieldBC: AxisLF is only valid for rho, v_z, e, and H_z "
                                                         "(fields with a defined half-stencil LF update).  '") +
                                             kCondNames[k] +
                                             "' has no AxisLF formula.  Use { dirichlet: 0.0 } for symmetry "
                                             "components that must be zero.");
            }
            g.has_axis = axis_main ? 1 : 0;
            for (int k = 0; k < 8; ++k) {
                MpdCond& c = g.c[k];
                c.value = fc[k]->value;
                switch (fc[k]->type) {
                    case mpd::FieldCondType::Neumann: c.type = MPD_NEUMANN; break;
                    case mpd::FieldCondType::Dirichlet: c.type = MPD_DIRICHLET; break;
                    case mpd::FieldCondType::AxisLF: c.type = MPD_AXIS_LF; break;
                    case mpd::FieldCondType::Expression: {
                        c.type = MPD_EXPRESSION;
                        g.has_expr = 1;
                        mpd::CompiledExpr e;
                        try {
                            e = mpd::CompileExpr(fc[k]->expr_str, names);
                        } catch (const std::exception& ex) {
                            throw std::runtime_error(std::string("PerFieldBC: failed to compile BC expression for field '") +
                                                     kCondNames[k] + "':\n  expression : " + fc[k]->expr_str +
                                                     "\n  error(s)   : " + ex.what());
                        }
                        for (int sym : e.symbols_used) {
                            if (sym >= MPD_BS_RHO && sym - MPD_BS_RHO >= k) {
                                throw std::runtime_error(
                                    std::string("PerFieldBC: the expression for '") + kCondNames[k] + "' reads '" +
                                    kBcSymNames[sym] +
                                    "', which is not evaluated before it in the order rho, v_z, v_r, v_phi, H_z, "
                                    "H_r, H_phi, e.  Such a forward reference reads the previous cell's value in the "
                                    "reference; it is not supported here.  Use the '_nb' symbol or reorder.");
                            }
                        }
                        if (ncode + (int)e.code.size() > MPD_MAX_CODE)
                            throw std::runtime_error("boundary_conditions: expressions too long in total");
                        c.code_off = ncode;
                        c.code_len = (int)e.code.size();
                        bc.max_stack = std::max(bc.max_stack, e.max_stack);
                        for (const MpdInstr& in : e.code) bc.code[ncode++] = in;
                        break;
                    }
                }
            }
        }
    }
    bc.ncode = ncode;
}

// The host half of solver construction (src/main.cpp:28-185): config,
// geometry, grid, expression IC, conservative state and the BC program.
struct HostSetup {
    mpd::SimConfig cfg;
    mpd::HostGrid grid;
    mpd::HostFields f;
    MpdBcProgram bc;
    mpd::StencilCfg scfg;

    // Per-row wall data of a grid computed by the caller (rows = L_max + 2
    // entries each); when given, the config's geometry section is not used.
    struct RowGeometry {
        const double* r_in;
        const double* r_out;
        const double* dr_in_dz;
        const double* dr_out_dz;
    };

    HostSetup(const std::string& src, bool is_file = false, const RowGeometry* rg = nullptr) {
        mpd::RegisterAllGeometries();
        if (is_file) cfg.Load(src);
        else cfg.LoadFromString(src);
        if (cfg.L_max < 3 || cfg.M_max < 2) throw std::runtime_error("grid: L_max must be >= 3 and M_max >= 2");
        if (cfg.convergence_threshold > 0.0 && cfg.check_frequency <= 0)
            throw std::runtime_error("convergence: check_frequency must be positive");
        if (cfg.gpu.kernel == "simple") scfg.kind = mpd::STENCIL_SIMPLE;
        else if (cfg.gpu.kernel == "tiled") scfg.kind = mpd::STENCIL_TILED;
        else throw std::runtime_error("gpu.kernel must be 'tiled' or 'simple'");
        scfg.fast_div = cfg.gpu.fast_div;
        scfg.tile_rows = cfg.gpu.tile_rows;
        // Initial condition: VTK restart when initial_conditions.vtk_file is
        // set, the expression IC otherwise (src/main.cpp:117-136).
        std::unique_ptr<mpd::IInitialCondition> ic;
        if (!cfg.initial_conditions.vtk_file.empty())
            ic = std::make_unique<mpd::VtkIC>(cfg.initial_conditions.vtk_file, cfg.initial_conditions.vtk_exact_z);
        else
            ic = std::make_unique<mpd::ExpressionIC>(
                cfg.initial_conditions.has_params ? &cfg.initial_conditions.params : nullptr,
                cfg.rho_noise_amplitude, cfg.rho_noise_seed);
        if (rg) {
            const int rows = cfg.L_max + 2;
            const double* arr[4] = {rg->r_in, rg->r_out, rg->dr_in_dz, rg->dr_out_dz};
            for (const double* a : arr) {
                if (!a) throw ArgError("null grid row array");
                for (int l = 0; l < rows; ++l)
                    if (!std::isfinite(a[l])) throw ArgError("grid row arrays must be finite");
            }
            grid.BuildFromRows(cfg, rg->r_in, rg->r_out, rg->dr_in_dz, rg->dr_out_dz);
        } else {
            auto geom = mpd::GeometryRegistry::Instance().Create(cfg.geometry.type, cfg.geometry.params);
            grid.Build(cfg, *geom);
        }
        CompileBc(cfg, bc);
        f.Allocate(grid.rows, grid.cols);
        ic->Apply(f, grid, cfg, 0);
        f.InitConservative(grid);
    }
};

}  // namespace mpd_detail

using namespace mpd_detail;

struct mpd_solver {
    mpd::SimConfig cfg;
    mpd::HostGrid grid;
    int rows = 0, cols = 0, pitch = 0;
    int64_t fstride = 0;
    double* state[2] = {nullptr, nullptr};
    int cur = 0;
    double* rowbuf = nullptr;
    double* phys_ic = nullptr;
    double* prev = nullptr;
    double* partial = nullptr;
    double* rowsum = nullptr;
    double* scratch = nullptr;
    double* geo = nullptr;
    double* ringgeo = nullptr;  // boundary geometry table (RingGeo layout, lf_device.h)
    uint64_t* bc = nullptr;  // the BC program's image (BcImage layout)
    MpdLoopState* st = nullptr;
    unsigned long long* umax = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;

    bool step1 = false;
    int speed_src = 0;  // 0: unknown, 1: st->speed_bits (explicit steps), 2: cached value
    double speed = 0.0;

    double t = 0.0, dt = 0.0, change = -1.0;
    int step = 0, checked_step = 0;
    bool converged = false;

    mpd::StencilCfg scfg;
    int steps_per_graph = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_K = 0;
    std::vector<double> ic_prev_host;  // IC snapshot for a lazily allocated prev

    // A solver inside a batch runs on the batch's stream and keeps its loop
    // state in the batch's array of loop states (neither owned here).
    bool own_stream = true, own_st = true;
    bool in_batch = false;  // owned by an mpd_batch (mpd_destroy leaves it alone)
    int bc_words = 0;       // LDS image size of the BC program (8-byte words)
    int bc_area = 0;        // per-lane work-area slots of the boundary kernel

    ~mpd_solver() {
        for (auto& g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        double* bufs[] = {state[0], state[1], rowbuf, phys_ic, prev, partial, rowsum, scratch, geo, ringgeo};
        for (double* b : bufs)
            if (b) (void)hipFree(b);
        if (bc) (void)hipFree(bc);
        if (st && own_st) (void)hipFree(st);
        if (umax) (void)hipFree(umax);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }

    StepArgs Args(int in) const {
        StepArgs a{};
        a.L = grid.L;
        a.Mn = grid.Mn;
        a.pitch = pitch;
        a.fstride = fstride;
        a.u_in = state[in];
        a.u_out = state[1 - in];
        a.rowbuf = rowbuf;
        a.phys_in = phys_ic;
        a.rin = geo;
        a.rout = geo + rows;
        a.drin = geo + 2 * rows;
        a.drout = geo + 3 * rows;
        a.dr = geo + 4 * rows;
        a.dr2inv = geo + 5 * rows;
        a.gamma = cfg.gamma;
        a.gm1 = cfg.gamma - 1.0;
        a.beta = cfg.beta;
        a.H_z0 = cfg.H_z0;
        a.r_0 = grid.r_0;
        a.dz = cfg.dz;
        a.dy = cfg.dy;
        a.dz2 = 2 * cfg.dz;
        const double dx = std::min(cfg.dz, cfg.dy);
        a.cfl_dx = cfg.cfl_number * dx;
        a.growth = cfg.dt_growth_factor;
        a.dt_min = cfg.dt_min;
        a.dt_max = cfg.dt_max;
        a.T = cfg.T;
        a.adaptive = cfg.adaptive_dt ? 1 : 0;
        a.step1 = step1 ? 1 : 0;
        a.bc_words = bc_words;
        a.bc_area = bc_area;
        a.st = st;
        a.bc = bc;
        a.ringgeo = ringgeo;
        return a;
    }

    // The boundary geometry table: for every boundary cell in compact order,
    // r and r_z of the cell, its row and column neighbours and the diagonal,
    // dr and z of its row and of the row neighbour's (the host's HostGrid
    // values, bit-identical to the device's cell_r / cell_rz).
    void UploadRingGeo() {
        const int L = grid.L, Mn = grid.Mn;
        const int n = mpd::ring_lanes_compact(L, Mn);
        std::vector<double> t((size_t)mpd::RG_N * n);
        for (int c = 0; c < n; ++c) {
            int l = 0, m = 0;
            mpd::ring_cell_compact(c, L, Mn, l, m);
            const int lr = (l == 1) ? 2 : (l == L ? L - 1 : l);
            const int mc = (m == 1) ? 2 : (m == Mn ? Mn - 1 : m);
            const double v[mpd::RG_N] = {grid.r(l, m),   grid.r_z(l, m),   grid.dr[l],      (double)(l - 1) * cfg.dz,
                                         grid.r(lr, m),  grid.r_z(lr, m),  grid.dr[lr],     (double)(lr - 1) * cfg.dz,
                                         grid.r(l, mc),  grid.r_z(l, mc),  grid.r(lr, mc),  grid.r_z(lr, mc)};
            for (int e = 0; e < mpd::RG_N; ++e) t[(size_t)e * n + c] = v[e];
        }
        ringgeo = Alloc(t.size());
        HIPCHK(hipMemcpy(ringgeo, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }

    double* Alloc(size_t ndouble) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, ndouble * sizeof(double)));
        return static_cast<double*>(p);
    }

    // Pitched field buffers: nrow rows of `pitch` doubles in the strip layout
    // of lf_launch.h (element (row, m) at p[row * pitch + phys_col(m)]).
    double* AllocPitched(size_t nrow) { return Alloc(nrow * (size_t)pitch); }
    void ZeroPitched(double* p, size_t nrow) { HIPCHK(hipMemset(p, 0, nrow * (size_t)pitch * sizeof(double))); }

    // Columns m_lo .. m_hi of one row, split at the strip boundaries: calls
    // f(m, phys_col(m), n) for each run of n consecutive columns.
    template <class F>
    static void ForRuns(int m_lo, int m_hi, F f) {
        for (int m = m_lo; m <= m_hi;) {
            const int end = std::min(m_hi, (m / mpd::kStripCols) * mpd::kStripCols + mpd::kStripCols - 1);
            f(m, mpd::phys_col(m), end - m + 1);
            m = end + 1;
        }
    }

    // Copy nf host arrays (rows x cols) into a pitched device buffer with the
    // ghost ring zeroed.
    void UploadFields(double* dst, const double* const* src, int nf) {
        std::vector<double> stage((size_t)rows * pitch, 0.0);
        for (int k = 0; k < nf; ++k) {
            std::fill(stage.begin(), stage.end(), 0.0);
            for (int l = 1; l <= grid.L; ++l) {
                const double* row = src[k] + (size_t)l * cols;
                double* out = &stage[(size_t)l * pitch];
                ForRuns(1, grid.Mn, [&](int m, int pc, int n) { std::memcpy(out + pc, row + m, sizeof(double) * n); });
            }
            HIPCHK(hipMemcpy(dst + (int64_t)k * fstride, stage.data(), stage.size() * sizeof(double),
                             hipMemcpyHostToDevice));
        }
    }

    void DownloadFields(const double* srcdev, double* const* dst, int nf) {
        std::vector<double> stage((size_t)rows * pitch);
        for (int k = 0; k < nf; ++k) {
            if (!dst[k]) continue;
            HIPCHK(hipMemcpy(stage.data(), srcdev + (int64_t)k * fstride, stage.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
            for (int l = 0; l < rows; ++l) {
                const double* in = &stage[(size_t)l * pitch];
                double* row = dst[k] + (size_t)l * cols;
                ForRuns(0, cols - 1, [&](int m, int pc, int n) { std::memcpy(row + m, in + pc, sizeof(double) * n); });
            }
        }
    }

    // The physical arrays Fields holds after the last step (or the IC before
    // the first), rows x cols each.
    void DownloadPhys(double* const* phys) {
        if (step1) {
            DownloadFields(phys_ic, phys, 10);
            return;
        }
        double* tmp = AllocPitched((size_t)10 * rows);
        const hipError_t e = mpd::launch_phys_fields(Args(cur), tmp, stream);
        if (e == hipSuccess) {
            (void)hipStreamSynchronize(stream);
            try {
                DownloadFields(tmp, phys, 10);
            } catch (...) {
                (void)hipFree(tmp);
                throw;
            }
        }
        (void)hipFree(tmp);
        HIPCHK(e);
    }

    void Init(const std::string& src, bool is_file, const HostSetup::RowGeometry* rg = nullptr) {
        HostSetup h(src, is_file, rg);
        InitFrom(h);
    }

    // Device half of the construction from a finished host setup.
    void InitFrom(HostSetup& h, hipStream_t ext_stream = nullptr, MpdLoopState* ext_st = nullptr) {
        cfg = h.cfg;
        grid = h.grid;
        scfg = h.scfg;
        steps_per_graph = cfg.gpu.steps_per_graph;
        rows = grid.rows;
        cols = grid.cols;
        pitch = mpd::layout_pitch(cols);
        fstride = (int64_t)rows * pitch;
        mpd::HostFields& f = h.f;
        const MpdBcProgram& hbc = h.bc;

        HIPCHK(hipGetDevice(&device));
        if (ext_stream) {
            stream = ext_stream;
            own_stream = false;
        } else {
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        }
        state[0] = AllocPitched((size_t)8 * rows);
        state[1] = AllocPitched((size_t)8 * rows);
        ZeroPitched(state[0], (size_t)8 * rows);
        ZeroPitched(state[1], (size_t)8 * rows);
        rowbuf = AllocPitched(16);
        ZeroPitched(rowbuf, 16);
        partial = Alloc(2 * 1024);
        rowsum = Alloc((size_t)7 * rows);
        scratch = Alloc(64);
        geo = Alloc((size_t)6 * rows);
        std::vector<double> g((size_t)6 * rows);
        for (int l = 0; l < rows; ++l) {
            g[l] = grid.rin[l];
            g[rows + l] = grid.rout[l];
            g[2 * rows + l] = grid.drin[l];
            g[3 * rows + l] = grid.drout[l];
            g[4 * rows + l] = grid.dr[l];
            g[5 * rows + l] = 0.0;  // div_recip(2 dr), filled on the device below
        }
        HIPCHK(hipMemcpy(geo, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(mpd::launch_recip_table(geo + 4 * rows, geo + 5 * rows, rows, stream));
        HIPCHK(hipStreamSynchronize(stream));
        bc_words = mpd::bc_image_words(hbc);
        {
            std::vector<uint64_t> image((size_t)bc_words);
            mpd::build_bc_image(hbc, image.data());
            HIPCHK(hipMalloc((void**)&bc, image.size() * sizeof(uint64_t)));
            HIPCHK(hipMemcpy(bc, image.data(), image.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        }
        bc_area = mpd::bc_area_slots(hbc);
        UploadRingGeo();
        if (ext_st) {
            st = ext_st;
            own_st = false;
        } else {
            HIPCHK(hipMalloc((void**)&st, sizeof(MpdLoopState)));
        }
        HIPCHK(hipMemset(st, 0, sizeof(MpdLoopState)));
        HIPCHK(hipMalloc((void**)&umax, sizeof(unsigned long long)));

        const double* u0p[8];
        for (int k = 0; k < 8; ++k) u0p[k] = f.u0[k].data();
        const double* php[10];
        for (int k = 0; k < 10; ++k) php[k] = f.phys[k].data();
        UploadState(u0p, php);
        // Fields::SavePrev at init (src/main.cpp:150-152): snapshot of the IC.
        if (cfg.convergence_threshold > 0.0) EnsurePrevFromPhys(php);
        t = 0.0;
        dt = cfg.dt;
        step = 0;
    }

    void EnsurePrevFromPhys(const double* const* php) {
        if (!prev) prev = AllocPitched((size_t)7 * rows);
        const double* p7[7] = {php[mpd::F_RHO], php[mpd::F_VZ], php[mpd::F_VR], php[mpd::F_VPHI],
                               php[mpd::F_HZ], php[mpd::F_HR], php[mpd::F_HPHI]};
        UploadFields(prev, p7, 7);
    }

    // The loop variables of a fresh run (src/main.cpp:191-209): t = 0,
    // dt = time.dt, no steps taken, no convergence check yet.
    void ResetLoop() {
        t = 0.0;
        dt = cfg.dt;
        step = 0;
        checked_step = 0;
        converged = false;
        change = -1.0;
        speed_src = 0;
    }

    // Host form of split_div_safe (lf_device.h) over the interior of host
    // arrays u0 (rows x cols each), and the radius range it assumes.
    static bool HostExpOk(double v, int lo, int hi) {
        int e = 0;
        if (v == 0.0) return true;
        if (!std::isfinite(v)) return true;  // a non-finite state ends the loop anyway
        (void)std::frexp(v, &e);
        return e >= lo && e <= hi;
    }
    bool StateInSplitRange(const double* const* u0) const {
        for (int l = 1; l <= grid.L; ++l)
            for (int m = 1; m <= grid.Mn; ++m) {
                const size_t i = (size_t)l * cols + m;
                if (!HostExpOk(u0[0][i], -29, 31) || u0[0][i] == 0.0) return false;
                for (int k = 1; k < 8; ++k)
                    if (!HostExpOk(u0[k][i], -99, 61)) return false;
            }
        return true;
    }
    bool RadiiInSplitRange() const {
        for (int l = 1; l <= grid.L; ++l)
            for (int m = 1; m <= grid.Mn; ++m) {
                const double r = grid.r(l, m);
                if (!(r >= 0.00390625 && r <= 256.0)) return false;
            }
        return true;
    }
    // The split division is used only while every state stays in the range
    // where it equals IEEE division; from the step after one leaves it, the
    // solver runs with IEEE division (fast_div_step records when).
    int fast_div_step = -1;
    void SwitchToIeeeDivision() {
        if (!scfg.fast_div) return;
        scfg.fast_div = false;
        fast_div_step = step;
        for (auto& g : graph)
            if (g) {
                HIPCHK(hipGraphExecDestroy(g));
                g = nullptr;
            }
        graph_K = 0;
    }

    void UploadState(const double* const* u0, const double* const* phys) {
        HIPCHK(hipStreamSynchronize(stream));
        if (scfg.fast_div && !(RadiiInSplitRange() && StateInSplitRange(u0))) SwitchToIeeeDivision();
        UploadFields(state[cur], u0, 8);
        ZeroPitched(state[1 - cur], (size_t)8 * rows);
        if (phys) {
            if (!phys_ic) phys_ic = AllocPitched((size_t)10 * rows);
            UploadFields(phys_ic, phys, 10);
            step1 = true;
        } else {
            step1 = false;
        }
        speed_src = 0;
    }

    void Advance(double dt_step) {
        if (scfg.fast_div) {
            // The previous step's kernels flag a new state outside the split
            // division's range (range_bits): this step then divides exactly.
            uint32_t bits = 0;
            HIPCHK(hipMemcpyAsync(&bits, &st->range_bits, sizeof(bits), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (bits) {
                SwitchToIeeeDivision();
                HIPCHK(hipMemsetAsync(&st->range_bits, 0, sizeof(uint32_t), stream));
            }
        }
        HIPCHK(hipMemsetAsync(&st->speed_bits, 0, sizeof(uint64_t), stream));
        const StepArgs a = Args(cur);
        HIPCHK(mpd::launch_stencil(a, scfg, dt_step, 0, 1, stream));
        HIPCHK(mpd::launch_bc(a, dt_step, 0, 1, stream));
        cur ^= 1;
        step1 = false;
        speed_src = 1;
    }

    double MaxWaveSpeed() {
        if (speed_src == 2) return speed;
        unsigned long long bits = 0;
        if (speed_src == 1) {
            HIPCHK(hipMemcpyAsync(&bits, &st->speed_bits, sizeof(bits), hipMemcpyDeviceToHost, stream));
        } else {
            HIPCHK(hipMemsetAsync(umax, 0, sizeof(unsigned long long), stream));
            HIPCHK(mpd::launch_max_speed(Args(cur), umax, stream));
            HIPCHK(hipMemcpyAsync(&bits, umax, sizeof(bits), hipMemcpyDeviceToHost, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        double v;
        std::memcpy(&v, &bits, sizeof(v));
        speed = v;
        speed_src = 2;
        return v;
    }

    // Diagnostics::ComputeDt, src/diagnostics.cpp:49-65.
    double ComputeDt(double dt_current) {
        const double s = MaxWaveSpeed();
        const double dx = std::min(cfg.dz, cfg.dy);
        const double dt_cfl = cfg.cfl_number * dx / (s + 1.0e-10);
        const double dt_grown = dt_current * cfg.dt_growth_factor;
        return std::clamp(std::min(dt_cfl, dt_grown), cfg.dt_min, cfg.dt_max);
    }

    double SolutionChange(bool save) {
        if (!prev) {
            // No snapshot yet: take the current state as the snapshot first.
            prev = AllocPitched((size_t)7 * rows);
            HIPCHK(hipMemsetAsync(prev, 0, (size_t)7 * fstride * sizeof(double), stream));
            HIPCHK(mpd::launch_residual(Args(cur), prev, partial, 0, 1, 0.0, 1, stream));
        }
        HIPCHK(mpd::launch_residual(Args(cur), prev, partial, 0, 1, 0.0, save ? 1 : 0, stream));
        double v = 0;
        HIPCHK(hipMemcpyAsync(&v, &st->change, sizeof(v), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return v;
    }

    int ChooseK() const {
        if (steps_per_graph > 0) return std::max(2, steps_per_graph + (steps_per_graph & 1));
        const double n = (double)grid.L * grid.Mn;
        int k = (int)(4.0e8 / n);
        k = std::max(2, std::min(100, k));
        return k + (k & 1);
    }

    // One step of the time loop.  `rescan`: also the exact wave-speed rescan
    // kernel, which finishes the step when its candidate max could not be
    // proven (rare; it exits at once otherwise).  The graphs leave it out: a
    // step that needs it idles the rest of the graph and the host finishes it
    // (ResolveRescan) before it launches the remaining steps.
    void EnqueueLoopStep(int in, bool conv, bool rescan) {
        const StepArgs a = Args(in);
        HIPCHK(mpd::launch_stencil(a, scfg, 0.0, 1, 1, stream));
        HIPCHK(mpd::launch_bc(a, 0.0, 1, 1, stream));
        if (rescan) HIPCHK(mpd::launch_speed_fallback(a, stream));
        if (conv) {
            StepArgs r = Args(1 - in);
            r.step1 = 0;
            HIPCHK(mpd::launch_residual(r, prev, partial, 1, cfg.check_frequency, cfg.convergence_threshold, 1,
                                        stream));
        }
    }

    // Finishes the step that waits for its exact wave-speed rescan (loop
    // state `fallback`; cur is that step's input): the rescan with the step's
    // dt bookkeeping, then the step's convergence check.
    void ResolveRescan(bool conv) {
        const StepArgs a = Args(cur);
        HIPCHK(mpd::launch_speed_fallback(a, stream));
        if (conv) {
            StepArgs r = Args(1 - cur);
            r.step1 = 0;
            HIPCHK(mpd::launch_residual(r, prev, partial, 1, cfg.check_frequency, cfg.convergence_threshold, 1,
                                        stream));
        }
    }

    void CaptureGraphs(int K, bool conv) {
        for (auto& g : graph)
            if (g) {
                HIPCHK(hipGraphExecDestroy(g));
                g = nullptr;
            }
        for (int p = 0; p < 2; ++p) {
            hipGraph_t gr = nullptr;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < K; ++i) EnqueueLoopStep((p + i) & 1, conv, false);
            HIPCHK(hipStreamEndCapture(stream, &gr));
            HIPCHK(hipGraphInstantiate(&graph[p], gr, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(gr));
        }
        graph_K = K;
    }

    MpdLoopState ReadLoop() {
        MpdLoopState h;
        HIPCHK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return h;
    }

    // The device loop state that continues the host mirror's loop for at
    // most max_steps more steps (<= 0: no cap besides time.max_steps).
    MpdLoopState LoopSeed(int max_steps) const {
        long long cap = LLONG_MAX;
        if (cfg.max_steps > 0) cap = cfg.max_steps;
        if (max_steps > 0) cap = std::min(cap, (long long)step + max_steps);
        MpdLoopState h{};
        h.t = t;
        h.dt = dt;
        h.dt_use = (t + dt > cfg.T) ? cfg.T - t : dt;
        h.change = change;
        h.max_speed = (speed_src == 2) ? speed : 0.0;  // the stencil's candidate threshold
        h.step = step;
        h.max_steps = (cap == LLONG_MAX) ? 0 : (int)std::min(cap, (long long)INT_MAX);
        h.converged = converged ? 1 : 0;
        h.done = (!(t < cfg.T) || converged || (h.max_steps > 0 && step >= h.max_steps)) ? 1 : 0;
        h.checked_step = checked_step;
        return h;
    }

    // Seeds the device loop state from the host mirror; returns done.
    bool SeedLoop(int max_steps) {
        const MpdLoopState h = LoopSeed(max_steps);
        HIPCHK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return h.done != 0;
    }

    void Absorb(const MpdLoopState& h) {
        const int ran = h.step - step;
        if (ran & 1) cur ^= 1;
        if (ran > 0) {
            step1 = false;
            speed = h.max_speed;
            speed_src = 2;
        }
        t = h.t;
        dt = h.dt;
        step = h.step;
        change = h.change;
        converged = h.converged != 0;
        checked_step = h.checked_step;
    }

    // Tile height of the tiled stencil when the config leaves it open
    // (gpu.tile_rows 0).  Large grids: 8 rows (measured fastest at
    // 6400x2560).  Grids below kTuneCells: a small grid gives few tiles, so a
    // step is the latency of one tile's march; the heights {1, 2, 4, 8} are
    // timed once on the current state (explicit dt: the launches write only
    // the next-state buffer and the row buffer, which a step rewrites before
    // reading, and the running max, which the loop seed resets) and the
    // fastest kept.
    static constexpr double kTuneCells = 4.0e6;
    bool rows_tuned = false;
    void TuneTileRows() {
        rows_tuned = true;
        if (scfg.kind != mpd::STENCIL_TILED || scfg.tile_rows > 0) return;
        if ((double)grid.L * grid.Mn >= kTuneCells) return;
        const StepArgs a = Args(cur);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int best = 8;
        float best_ms = 1e30f;
        for (int h : {8, 4, 2, 1}) {
            if (h > grid.L) continue;
            mpd::StencilCfg c = scfg;
            c.tile_rows = h;
            StepArgs t = a;
            t.step1 = 0;
            HIPCHK(mpd::launch_stencil(t, c, 1.0e-9, 0, 1, stream));
            HIPCHK(hipEventRecord(e0, stream));
            for (int r = 0; r < 8; ++r) HIPCHK(mpd::launch_stencil(t, c, 1.0e-9, 0, 1, stream));
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = h;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        scfg.tile_rows = best;
    }

    // Steps this call may still take (max_steps <= 0: no cap of its own).
    static int Remaining(int max_steps, int taken) { return max_steps > 0 ? max_steps - taken : 0; }

    // After a step whose new state left the split division's range (the loop
    // idles behind range_stop): IEEE division from here, the loop re-seeded.
    // Returns whether the loop has ended.
    bool AfterRangeStop(const MpdLoopState& h, int max_steps, int step0) {
        SwitchToIeeeDivision();
        if (h.done) return true;
        if (max_steps > 0 && step - step0 >= max_steps) return true;
        return SeedLoop(Remaining(max_steps, step - step0));
    }

    void Run(int max_steps, mpd_run_result* out) {
        const int step0 = step;
        const bool conv = cfg.convergence_threshold > 0.0;
        if (conv && !prev) throw std::runtime_error("convergence snapshot missing");
        if (!rows_tuned) TuneTileRows();
        bool done = SeedLoop(max_steps);
        if (!done && step1) {
            EnqueueLoopStep(cur, conv, true);
            const MpdLoopState h = ReadLoop();
            Absorb(h);
            done = h.range_stop ? AfterRangeStop(h, max_steps, step0) : (h.done != 0);
        }
        while (!done) {
            const int K = ChooseK();
            if (!graph[0] || graph_K != K) CaptureGraphs(K, conv);
            HIPCHK(hipGraphLaunch(graph[cur], stream));
            MpdLoopState h = ReadLoop();
            Absorb(h);
            if (!h.done && h.fallback) {
                ResolveRescan(conv);
                h = ReadLoop();
                Absorb(h);
            }
            done = h.range_stop ? AfterRangeStop(h, max_steps, step0) : (h.done != 0);
        }
        if (out) {
            out->t = t;
            out->dt = dt;
            out->change = change;
            out->steps = step - step0;
            out->total_steps = step;
            out->converged = converged ? 1 : 0;
            // The loop's own end (src/main.cpp:219), not this call's step cap.
            out->done = (!(t < cfg.T) || converged || (cfg.max_steps > 0 && step >= cfg.max_steps)) ? 1 : 0;
            out->fast_div = scfg.fast_div ? 1 : 0;
        }
    }

    void TimeSteps(int steps, int warmup, double* total_ms, double* stencil_ms, double* bc_ms) {
        const bool conv = cfg.convergence_threshold > 0.0;
        if (warmup > 0) Run(warmup, nullptr);
        if (step1) Run(1, nullptr);
        const int K = ChooseK();
        if (!graph[0] || graph_K != K) CaptureGraphs(K, conv);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        // Whole loop, graphs back to back, then the steps short of a whole
        // graph one by one (each with its rescan kernel), in a segment of
        // their own: a graph step that waits for its exact wave-speed rescan
        // idles the rest of the segment, the host runs the rescan and
        // launches the remaining steps (all of it timed).
        if (SeedLoop(steps)) throw std::runtime_error("time loop already finished (t >= T or converged)");
        const int target = step + steps;
        double ms = 0.0;
        for (;;) {
            const int todo = target - step;
            HIPCHK(hipEventRecord(e0, stream));
            if (todo >= K) {
                for (int i = 0; i < todo / K; ++i) HIPCHK(hipGraphLaunch(graph[cur], stream));
            } else {
                for (int i = 0; i < todo; ++i) EnqueueLoopStep((cur + i) & 1, conv, true);
            }
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float part = 0.f;
            HIPCHK(hipEventElapsedTime(&part, e0, e1));
            ms += part;
            const int before = step;
            MpdLoopState h = ReadLoop();
            Absorb(h);
            if (h.done) break;
            if (!h.fallback) {
                if (step == before) throw std::runtime_error("time loop made no progress");
                continue;
            }
            HIPCHK(hipEventRecord(e0, stream));
            ResolveRescan(conv);
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&part, e0, e1));
            ms += part;
            h = ReadLoop();
            Absorb(h);
            if (h.done) break;
        }
        if (total_ms) *total_ms = ms;
        // Per-kernel averages: each launch bracketed by events.
        if (stencil_ms || bc_ms) {
            const int n = std::max(1, std::min(steps, 200));
            SeedLoop(n);
            std::vector<hipEvent_t> ev(3 * n + 1);
            for (auto& e : ev) HIPCHK(hipEventCreate(&e));
            int in = cur;
            HIPCHK(hipEventRecord(ev[0], stream));
            for (int i = 0; i < n; ++i) {
                const StepArgs a = Args(in);
                HIPCHK(mpd::launch_stencil(a, scfg, 0.0, 1, 1, stream));
                HIPCHK(hipEventRecord(ev[3 * i + 1], stream));
                HIPCHK(mpd::launch_bc(a, 0.0, 1, 1, stream));
                HIPCHK(hipEventRecord(ev[3 * i + 2], stream));
                HIPCHK(mpd::launch_speed_fallback(a, stream));
                if (conv)
                    HIPCHK(mpd::launch_residual(Args(1 - in), prev, partial, 1, cfg.check_frequency,
                                                cfg.convergence_threshold, 1, stream));
                HIPCHK(hipEventRecord(ev[3 * i + 3], stream));
                in ^= 1;
            }
            HIPCHK(hipEventSynchronize(ev[3 * n]));
            double s1 = 0, s2 = 0;
            for (int i = 0; i < n; ++i) {
                float a = 0, b = 0;
                HIPCHK(hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]));
                HIPCHK(hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]));
                s1 += a;
                s2 += b;
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
            Absorb(ReadLoop());
            if (stencil_ms) *stencil_ms = s1 / n;
            if (bc_ms) *bc_ms = s2 / n;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

namespace mpd_detail {

template <class F>
int Guard(F&& fn) {
    try {
        fn();
        return MPD_OK;
    } catch (const DeviceError& e) {
        g_last_error = e.what();
        return MPD_E_DEVICE;
    } catch (const ArgError& e) {
        g_last_error = e.what();
        return MPD_E_ARG;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return MPD_E_DEVICE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return MPD_E_CONFIG;
    }
}

inline void Need(const void* p) {
    if (!p) throw ArgError("null argument");
}

}  // namespace mpd_detail
