// ref_harness.cpp -- drives the reference's own compiled solver code (test
// infrastructure only; built into oracle/_ref/ by oracle/Makefile when the
// reference sources are present, never shipped or called by the product).
//
// Linked reference translation units, compiled from the reference tree as is:
//   src/solver.cpp src/face_bc.cpp src/fields.cpp src/array2d.cpp src/grid.cpp
//   src/mpi_manager.cpp src/diagnostics.cpp src/integrals.cpp src/verification.cpp
// Not built: src/config.cpp (needs yaml-cpp) and src/bcs/per_field_bc.cpp
// (needs exprtk).  The harness therefore configures no boundary segments, so
// FaceBC::FromConfig never constructs a PerFieldBC (the one symbol the link
// leaves unresolved), and it sets SimConfig's fields directly instead of
// SimConfig::Load; SimConfig's constructor calls Init(), defined below as an
// empty function because the harness assigns dz = 1/L_max and dy = 1/M_max
// itself (src/config.cpp:15-18).
//
// Usage: ref_harness L_max M_max geom p0 p1 p2 p3 p4 gamma steps dt adaptive cfl
//                    growth dt_min dt_max in.bin out.bin
// in.bin : 10 physical arrays (rho v_z v_r v_phi e p P H_z H_r H_phi) then the
//          8 u0 arrays, each (L_max+2) x (M_max+3) doubles, row-major.
// out.bin: see the write order at the end of main().
//
// Timing mode: ref_harness time L_max M_max steps builds the memory-bound
// channel workload in the process (converging-diverging walls r_in = 0.2 +
// 0.05 sin(pi z), r_out = 0.8 - 0.1 z^2; rho = 1 with +-1e-3 SplitMix64
// noise of seed 0, v_z = 0.1, H_phi = 0.2 / r, p = 0.5; gamma 5/3, adaptive
// CFL 0.4), runs one untimed step and then `steps` steps of the time loop
// (Solver::Advance + Diagnostics::ComputeDt), and prints the loop's wall
// time: the reference's own step timed as the CPU baseline.  No boundary
// segments (see above): the faces' O(L + M) work per step is left out.
#include <sched.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "config.hpp"
#include "diagnostics.hpp"
#include "fields.hpp"
#include "grid.hpp"
#include "igeometry.hpp"
#include "integrals.hpp"
#include "mpi_manager.hpp"
#include "solver.hpp"
#include "verification.hpp"

void SimConfig::Init() {}

namespace {

// The harness's geometry plug-in (the IGeometry interface of
// include/igeometry.hpp:16-34): straight coaxial channel (geom 0) or the
// converging-diverging channel r_in = p0 + p1 sin(pi z), r_out = p2 - p3 z^2
// (geom 1).
class HarnessGeometry : public IGeometry {
public:
    HarnessGeometry(int type, const double* p) : type_(type) {
        for (int i = 0; i < 5; ++i) p_[i] = p[i];
    }
    auto RInner(double z) const -> double override {
        return type_ == 0 ? p_[0] : p_[0] + p_[1] * std::sin(M_PI * z);
    }
    auto ROuter(double z) const -> double override {
        return type_ == 0 ? p_[1] : p_[2] - p_[3] * z * z;
    }
    auto DrInnerDz(double z) const -> double override {
        return type_ == 0 ? 0.0 : p_[1] * M_PI * std::cos(M_PI * z);
    }
    auto DrOuterDz(double z) const -> double override {
        return type_ == 0 ? 0.0 : -2.0 * p_[3] * z;
    }
    auto Name() const -> std::string override { return "harness"; }

private:
    int type_;
    double p_[5];
};

void Put(std::FILE* f, const double* p, size_t n) {
    if (std::fwrite(p, sizeof(double), n, f) != n) {
        std::fprintf(stderr, "ref_harness: short write\n");
        std::exit(3);
    }
}

void PutArray(std::FILE* f, const Array2D& a) { Put(f, a.Flat(), a.Size()); }

void PutScalar(std::FILE* f, double v) { Put(f, &v, 1); }

void GetArray(std::FILE* f, Array2D& a) {
    if (std::fread(a.Flat(), sizeof(double), a.Size(), f) != a.Size()) {
        std::fprintf(stderr, "ref_harness: short read\n");
        std::exit(3);
    }
}

}  // namespace

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) of (seed, idx) as a uniform in [0, 1).
double NoiseU01(unsigned long long seed, unsigned long long idx) {
    unsigned long long z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int TimeMain(int argc, char** argv) {
    SimConfig cfg;
    cfg.L_max = std::atoi(argv[2]);
    cfg.M_max = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]);
    if (cfg.L_max < 2 || cfg.M_max < 2 || steps < 1) return 2;
    cfg.gamma = 1.6666666666666667;
    cfg.dt = 1.0e-4;
    cfg.adaptive_dt = true;
    cfg.cfl_number = 0.4;
    cfg.dt_growth_factor = 1.1;
    cfg.dt_min = 1.0e-12;
    cfg.dt_max = 1.0e-2;
    cfg.dz = 1.0 / cfg.L_max;
    cfg.dy = 1.0 / cfg.M_max;
    cfg.mpi_dims_l = 1;
    cfg.mpi_dims_m = 1;
    cfg.convergence_threshold = 1.0;
    int ac = 1;
    MPIManager mpi(ac, argv, cfg);
    const double gp[5] = {0.2, 0.05, 0.8, 0.1, 0.0};
    HarnessGeometry geom(1, gp);
    Grid grid(cfg, mpi.local_L_with_ghosts, mpi.l_start, mpi.local_M_with_ghosts, mpi.m_start, geom);
    Fields f(mpi.local_L_with_ghosts, mpi.local_M_with_ghosts, true);
    const int Mn = cfg.M_max + 1;
    const double gm1 = cfg.gamma - 1.0;
    for (int l = 1; l <= cfg.L_max; ++l) {
        for (int m = 1; m <= Mn; ++m) {
            const double r = grid.r[l][m];
            const double u = NoiseU01(0, (unsigned long long)(l - 1) * Mn + (m - 1));
            const double rho = 1.0 * (1.0 + 1.0e-3 * (2.0 * u - 1.0));
            f.rho[l][m] = rho;
            f.v_z[l][m] = 0.1;
            f.H_phi[l][m] = 0.2 / r;
            f.e[l][m] = 0.5 / (gm1 * rho);
            f.p[l][m] = gm1 * rho * f.e[l][m];
            f.P[l][m] = f.p[l][m] + 0.5 * (f.H_phi[l][m] * f.H_phi[l][m]);
        }
    }
    f.InitConservative(grid);
    Solver solver(cfg, mpi, grid, f);
    double dt = cfg.dt;
    solver.Advance(dt);
    dt = Diagnostics::ComputeDt(f, cfg, mpi.local_L, mpi.local_M, mpi, dt);
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        solver.Advance(dt);
        dt = Diagnostics::ComputeDt(f, cfg, mpi.local_L, mpi.local_M, mpi, dt);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    cpu_set_t set;
    CPU_ZERO(&set);
    const int aff = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : -1;
    std::printf("loop_seconds %.9f\nsteps %d\ndt %.17g\naffinity_cpus %d\n", secs, steps, dt, aff);
    (void)argc;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 5 && std::string(argv[1]) == "time") return TimeMain(argc, argv);
    if (argc != 19) {
        std::fprintf(stderr, "usage: ref_harness L M geom p0 p1 p2 p3 p4 gamma steps dt "
                             "adaptive cfl growth dt_min dt_max in out\n");
        return 2;
    }
    SimConfig cfg;
    cfg.L_max = std::atoi(argv[1]);
    cfg.M_max = std::atoi(argv[2]);
    const int geom_type = std::atoi(argv[3]);
    double gp[5];
    for (int i = 0; i < 5; ++i) gp[i] = std::strtod(argv[4 + i], nullptr);
    cfg.gamma = std::strtod(argv[9], nullptr);
    const int steps = std::atoi(argv[10]);
    cfg.dt = std::strtod(argv[11], nullptr);
    cfg.adaptive_dt = std::atoi(argv[12]) != 0;
    cfg.cfl_number = std::strtod(argv[13], nullptr);
    cfg.dt_growth_factor = std::strtod(argv[14], nullptr);
    cfg.dt_min = std::strtod(argv[15], nullptr);
    cfg.dt_max = std::strtod(argv[16], nullptr);
    cfg.dz = 1.0 / cfg.L_max;
    cfg.dy = 1.0 / cfg.M_max;
    cfg.mpi_dims_l = 1;
    cfg.mpi_dims_m = 1;
    cfg.convergence_threshold = 1.0;

    int ac = 1;
    char** av = argv;
    MPIManager mpi(ac, av, cfg);
    HarnessGeometry geom(geom_type, gp);
    Grid grid(cfg, mpi.local_L_with_ghosts, mpi.l_start, mpi.local_M_with_ghosts, mpi.m_start,
              geom);
    Fields f(mpi.local_L_with_ghosts, mpi.local_M_with_ghosts, true);

    std::FILE* in = std::fopen(argv[17], "rb");
    if (!in) {
        std::fprintf(stderr, "ref_harness: cannot open %s\n", argv[17]);
        return 3;
    }
    Array2D* phys[10] = {&f.rho, &f.v_z, &f.v_r, &f.v_phi, &f.e, &f.p, &f.P, &f.H_z, &f.H_r, &f.H_phi};
    Array2D* u0[8] = {&f.u0_1, &f.u0_2, &f.u0_3, &f.u0_4, &f.u0_5, &f.u0_6, &f.u0_7, &f.u0_8};
    for (auto* a : phys) GetArray(in, *a);
    for (auto* a : u0) GetArray(in, *a);
    std::fclose(in);

    std::FILE* out = std::fopen(argv[18], "wb");
    if (!out) return 3;

    // Grid (src/grid.cpp:20-42).
    PutArray(out, grid.r);
    PutArray(out, grid.r_z);
    Put(out, grid.R.data(), grid.R.size());
    Put(out, grid.dr.data(), grid.dr.size());
    PutScalar(out, grid.r_0);

    // Diagnostics of the input state.
    f.SavePrev();
    PutScalar(out, Diagnostics::MaxWaveSpeed(f, cfg, mpi.local_L, mpi.local_M, mpi));
    const auto q0 = Verification::ComputeIntegrals(f, grid, cfg, mpi);
    const double q0v[7] = {q0.mass, q0.momentum_z, q0.momentum_r, q0.momentum_phi,
                           q0.energy, q0.mag_flux_z, q0.mag_flux_phi};
    Put(out, q0v, 7);

    // Time loop of src/main.cpp:220-231 without the T truncation.
    Solver solver(cfg, mpi, grid, f);
    double dt = cfg.dt;
    std::vector<double> dts;
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        solver.Advance(dt);
        dts.push_back(dt);
        if (cfg.adaptive_dt) {
            dt = Diagnostics::ComputeDt(f, cfg, mpi.local_L, mpi.local_M, mpi, dt);
        }
    }
    // Wall time of the loop alone (the CPU-baseline calibration reads it).
    std::fprintf(stderr, "loop_seconds %.9f\n",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());

    for (auto* a : u0) PutArray(out, *a);
    for (auto* a : phys) PutArray(out, *a);
    Put(out, dts.data(), dts.size());
    PutScalar(out, dt);
    PutScalar(out, Diagnostics::SolutionChange(f, mpi.local_L, mpi.local_M));
    PutScalar(out, Diagnostics::MaxWaveSpeed(f, cfg, mpi.local_L, mpi.local_M, mpi));
    const auto q = Verification::ComputeIntegrals(f, grid, cfg, mpi);
    const double qv[7] = {q.mass, q.momentum_z, q.momentum_r, q.momentum_phi,
                          q.energy, q.mag_flux_z, q.mag_flux_phi};
    Put(out, qv, 7);
    PutScalar(out, Diagnostics::GetMassFlux(f, grid, cfg, mpi));
    PutScalar(out, Diagnostics::GetThrust(f, grid, cfg, mpi));
    std::fclose(out);
    return 0;
}
