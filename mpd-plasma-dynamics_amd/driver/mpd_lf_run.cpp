// mpd_lf_run.cpp -- the simulation driver over the C ABI: the reference's
// main (src/main.cpp:22-317) for one GPU holding the whole domain.  Same
// command line (mpd_lf_run [config.yaml]), same log lines, same run
// directory and VTK frames; the time loop runs on the device between the
// steps where the driver has something to print or write.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <filesystem>
#include <stdexcept>
#include <string>

#include "../../include/mpd_lf.h"
#include "../host/config.h"

namespace {

void Check(int rc) {
    if (rc != MPD_OK) throw std::runtime_error(mpd_last_error());
}

// Verification::PrintIntegralsHeader / PrintIntegrals, src/verification.cpp:49-70.
void PrintIntegralsHeader() {
    std::printf("\n[Verification] Conserved integrals\n");
    std::printf("  %-8s  %-12s  %-14s  %-14s  %-14s  %-14s  %-14s  %-14s  %-14s\n", "step", "t", "mass", "mom_z",
                "mom_r", "mom_phi", "energy", "mag_z", "mag_phi");
    std::printf("  %-8s  %-12s  %-14s  %-14s  %-14s  %-14s  %-14s  %-14s  %-14s\n", "--------", "------------",
                "--------------", "--------------", "--------------", "--------------", "--------------",
                "--------------", "--------------");
}

void PrintIntegrals(const double* q, int step, double t) {
    std::printf("  %-8d  %-12.5f  %-14.6e  %-14.6e  %-14.6e  %-14.6e  %-14.6e  %-14.6e  %-14.6e\n", step, t, q[0],
                q[1], q[2], q[3], q[4], q[5], q[6]);
}

// Verification::ReportDrift, src/verification.cpp:72-115.
void ReportDrift(const double* ref, const double* cur, int step, double t) {
    const double eps = 1.0e-30;
    const char* const names[7] = {"mass", "mom_z", "mom_r", "mom_phi", "energy", "mag_z", "mag_phi"};
    double scale = 0.0;
    for (int i = 0; i < 7; ++i) scale = std::max(scale, std::abs(ref[i]));
    if (scale < 1e-30)
        for (int i = 0; i < 7; ++i) scale = std::max(scale, std::abs(cur[i]));
    if (scale < 1e-30) scale = 1.0;
    double max_drift = 0.0;
    for (int i = 0; i < 7; ++i) max_drift = std::max(max_drift, std::abs(cur[i] - ref[i]) / scale);
    std::printf("[Verification] Drift from step-0 reference at step %d (t=%.5f)\n", step, t);
    std::printf("  %-10s  %-18s  %-18s  %-12s\n", "quantity", "reference", "current", "rel_drift");
    std::printf("  %-10s  %-18s  %-18s  %-12s\n", "----------", "------------------", "------------------",
                "------------");
    for (int i = 0; i < 7; ++i) {
        const double drift = std::abs(cur[i] - ref[i]) / (std::abs(ref[i]) + eps);
        std::printf("  %-10s  %+18.8e  %+18.8e  %12.4e\n", names[i], ref[i], cur[i], drift);
    }
    std::printf("  max_drift = %.4e\n\n", max_drift);
}

// IOManager's run directory, src/io_manager.cpp:27-59.
std::string MakeRunDir(const mpd::SimConfig& cfg) {
    using namespace std::chrono;
    const auto now_tp = system_clock::now();
    const auto ms = duration_cast<milliseconds>(now_tp.time_since_epoch()) % 1000;
    const std::time_t tt = system_clock::to_time_t(now_tp);
    struct tm tm_info {};
    localtime_r(&tt, &tm_info);
    char ts[64];
    std::snprintf(ts, sizeof(ts), "%02d-%02d-%04d_%02d:%02d:%02d:%03d", tm_info.tm_mday, tm_info.tm_mon + 1,
                  tm_info.tm_year + 1900, tm_info.tm_hour, tm_info.tm_min, tm_info.tm_sec, (int)ms.count());
    const std::string dir = cfg.output_dir + "/" + cfg.run_name + "_" + ts;
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) throw std::runtime_error("IOManager: cannot create run directory '" + dir + "': " + ec.message());
    std::printf("Run directory: %s\n", dir.c_str());
    return dir;
}

void WriteFrame(mpd_solver* s, const std::string& dir, int step) {
    char filename[1024];
    std::snprintf(filename, sizeof(filename), "%s/step_%04d.vtk", dir.c_str(), step);
    Check(mpd_write_vtk(s, filename));
}

int Main(int argc, char* argv[]) {
    const char* config_path = (argc > 1) ? argv[1] : "config.yaml";
    mpd::SimConfig cfg;
    cfg.Load(config_path);
    std::setvbuf(stdout, nullptr, _IONBF, 0);

    std::printf("Config file         : %s\n", config_path);
    std::printf("Grid                : %d x %d  (L x M)\n", cfg.L_max, cfg.M_max);
    std::printf("Initial dt / T end  : %.6e / %.4f\n", cfg.dt, cfg.T);
    std::printf("MPI ranks           : %d  (%d x %d Cartesian)\n", 1, 1, 1);
    std::printf("Geometry            : %s\n", cfg.geometry.type.c_str());
    if (cfg.adaptive_dt) {
        std::printf("Adaptive dt         : ON  (CFL=%.2f, growth=%.2f, dt_min=%.1e, dt_max=%.1e)\n", cfg.cfl_number,
                    cfg.dt_growth_factor, cfg.dt_min, cfg.dt_max);
    } else {
        std::printf("Adaptive dt         : OFF (fixed dt=%.6e)\n", cfg.dt);
    }
    if (cfg.convergence_threshold > 0.0)
        std::printf("Convergence check   : threshold = %.2e, every %d steps\n", cfg.convergence_threshold,
                    cfg.check_frequency);
    if (cfg.vtk_step > 0) std::printf("VTK output every    : %d steps\n", cfg.vtk_step);

    // The ghost-exchange round trip has nothing to exchange on one rank; it
    // passes as the reference's does without neighbours.
    std::printf("\n--- Layer-1 verification ---\n");
    std::printf("[Verification] Ghost exchange  : PASS\n");
    if (!cfg.initial_conditions.vtk_file.empty())
        std::printf("IC                  : VTK restart  '%s'\n", cfg.initial_conditions.vtk_file.c_str());
    else
        std::printf("IC                  : expression\n");

    mpd_solver* s = nullptr;
    Check(mpd_create_from_file(config_path, &s));
    struct Owner {
        mpd_solver* s;
        ~Owner() { mpd_destroy(s); }
    } owner{s};

    double sym = 0.0;
    Check(mpd_radial_symmetry(s, &sym));
    std::printf("[Verification] Radial symmetry : max |\xce\x94\xcf\x81|/<\xcf\x81> = %.3e  (%s)\n", sym,
                sym < 1.0e-10 ? "field is radially uniform" : "radial variation present");

    double ref[7];
    Check(mpd_integrals(s, ref));
    PrintIntegralsHeader();
    PrintIntegrals(ref, 0, 0.0);
    std::printf("----------------------------\n\n");

    const std::string run_dir = MakeRunDir(cfg);

    // Checkpoint cell of src/main.cpp:196-204: global (20, 40), local +1.
    const int check_l_global = 20, check_m_global = 40;
    const bool owns_checkpoint = check_l_global <= cfg.L_max - 1 && check_m_global <= cfg.M_max;
    std::printf("%-14s %-14s %-14s %-14s %-14s %-14s %-14s\n", "t", "dt", "rho", "v_z", "v_phi", "e", "H_phi");
    std::printf("%-14s %-14s %-14s %-14s %-14s %-14s %-14s\n", "--------------", "--------------", "--------------",
                "--------------", "--------------", "--------------", "--------------");
    if (cfg.vtk_step > 0) WriteFrame(s, run_dir, 0);

    const auto begin = std::chrono::steady_clock::now();
    int step_count = 0;
    double t = 0.0, dt = cfg.dt;
    const bool conv = cfg.convergence_threshold > 0.0;
    const double dx = std::min(cfg.dz, cfg.dy);
    for (;;) {
        // Next step at which the driver acts: a convergence report, a CFL
        // check, a VTK frame or a checkpoint line (src/main.cpp:233-281).
        auto next_multiple = [&](int k) { return k > 0 ? (step_count / k + 1) * k : 0x7fffffff; };
        int next = next_multiple(1000);
        if (conv) next = std::min(next, next_multiple(cfg.check_frequency));
        if (cfg.vtk_step > 0) next = std::min(next, next_multiple(cfg.vtk_step));
        mpd_run_result r{};
        Check(mpd_run(s, next - step_count, &r));
        step_count = r.total_steps;
        t = r.t;
        dt = r.dt;
        const bool converged = r.converged != 0;
        if (conv && r.steps > 0 && step_count % cfg.check_frequency == 0) {
            std::printf("Step %d, t=%.6f, dt=%.6e, relative change: %.6e\n", step_count, t, dt, r.change);
            if (converged) std::printf("Converged at t=%.6f after %d steps\n", t, step_count);
        }
        if (r.steps > 0 && step_count % 1000 == 0) {
            // Diagnostics::CheckCfl (src/diagnostics.cpp:114-125) runs every 100 steps and
            // prints only on multiples of 1000; nothing else depends on it.
            double speed = 0.0;
            Check(mpd_max_wave_speed(s, &speed));
            const double dt_max = cfg.cfl_number * dx / (speed + 1.0e-10);
            if (dt > dt_max)
                std::printf("WARNING [step %d]: dt=%.6e exceeds CFL limit dt_max=%.6e (max_speed=%.3f, CFL=%.2f)\n",
                            step_count, dt, dt_max, speed, dt / dt_max * cfg.cfl_number);
        }
        if (r.steps > 0 && cfg.vtk_step > 0 && step_count % cfg.vtk_step == 0) WriteFrame(s, run_dir, step_count);
        if (r.steps > 0 && step_count % 1000 == 0) {
            double v[10] = {0};
            if (owns_checkpoint) Check(mpd_probe(s, check_l_global + 1, check_m_global + 1, v));
            std::printf("%-14.6f %-14.6e %-14.6f %-14.6f %-14.6f %-14.6f %-14.6f\n", t, dt, v[0], v[1], v[3], v[4],
                        v[9]);
            double q[7];
            Check(mpd_integrals(s, q));
            PrintIntegrals(q, step_count, t);
        }
        if (r.done || r.steps == 0) break;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    std::printf("\nCalculation time : %.3f sec  (%d steps)\n", secs, step_count);
    WriteFrame(s, run_dir, step_count);
    std::printf("Final VTK written : %s/step_%04d.vtk\n", run_dir.c_str(), step_count);

    double fin[7];
    Check(mpd_integrals(s, fin));
    ReportDrift(ref, fin, step_count, t);

    double mass_flux = 0.0, thrust = 0.0;
    Check(mpd_outlet(s, &mass_flux, &thrust));
    std::printf("\n--- Outlet-plane diagnostics ---\n");
    std::printf("Mass flux : %.6e\n", mass_flux);
    std::printf("Thrust    : %.6e\n", thrust);
    std::printf("--------------------------------\n");
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    try {
        return Main(argc, argv);
    } catch (const std::exception& e) {
        std::fflush(stdout);
        std::fprintf(stderr, "mpd_lf_run: %s\n", e.what());
        return 1;
    }
}
