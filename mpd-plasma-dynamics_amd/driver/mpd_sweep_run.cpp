// mpd_sweep_run.cpp -- a parameter sweep over the batch C ABI: every
// config.yaml given on the command line is one case, the reference's whole
// run (src/main.cpp:22-317: its time loop, dt policy, T, step cap and
// convergence check) for that config, and all of them advance together on
// one GPU.  For each case the driver prints the loop's end state and the
// outlet diagnostics the reference prints last (src/main.cpp:306-314), and
// with --out writes the case's final VTK frame (IOManager::WriteFrame's
// format, src/io_manager.cpp:218-315).
//
//   mpd_sweep_run [--max-steps N] [--no-autotune] [--out DIR] a.yaml [b.yaml ...]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpd_lf.h"

namespace {

void Check(int rc) {
    if (rc != MPD_OK) throw std::runtime_error(mpd_last_error());
}

std::string ReadFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open config file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

int Usage() {
    std::fprintf(stderr, "usage: mpd_sweep_run [--max-steps N] [--no-autotune] [--out DIR] a.yaml [b.yaml ...]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    int max_steps = 0;
    bool autotune = true;
    std::string out_dir;
    std::vector<std::string> paths;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--max-steps" && i + 1 < argc) max_steps = std::atoi(argv[++i]);
        else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
        else if (a == "--no-autotune") autotune = false;
        else if (!a.empty() && a[0] == '-') return Usage();
        else paths.push_back(a);
    }
    if (paths.empty()) return Usage();
    try {
        std::vector<std::string> texts;
        for (const auto& p : paths) texts.push_back(ReadFile(p));
        std::vector<const char*> ptrs;
        for (const auto& t : texts) ptrs.push_back(t.c_str());
        const int n = (int)texts.size();
        mpd_batch* b = nullptr;
        Check(mpd_batch_create(ptrs.data(), n, &b));
        std::printf("[sweep] %d cases\n", n);
        if (autotune) {
            char report[4096];
            Check(mpd_batch_autotune(b, 5, report, (int)sizeof report));
            std::printf("[sweep] autotuner:\n%s", report);
        }
        std::vector<mpd_run_result> res(n);
        const auto t0 = std::chrono::steady_clock::now();
        if (max_steps > 0 || !autotune) {
            Check(mpd_batch_run(b, max_steps, res.data()));
        } else {
            // To the loops' own ends in slices of kSlice steps (the same bits
            // as one run): when the cases still running hold under a quarter
            // of the cells the tile heights were tuned for, tune them again
            // -- the tail of a few large cases wants shorter tiles.
            constexpr int kSlice = 20000;
            std::vector<long long> cells(n);
            long long tuned = 0;
            for (int i = 0; i < n; ++i) {
                mpd_solver* s = nullptr;
                Check(mpd_batch_instance(b, i, &s));
                int L = 0, M = 0, rows = 0, cols = 0;
                Check(mpd_dims(s, &L, &M, &rows, &cols));
                cells[i] = (long long)L * (M + 1);
                tuned += cells[i];
            }
            for (;;) {
                Check(mpd_batch_run(b, kSlice, res.data()));
                long long live = 0;
                for (int i = 0; i < n; ++i)
                    if (!res[i].done) live += cells[i];
                if (live == 0) break;
                if (live < tuned / 4) {
                    char report[4096];
                    Check(mpd_batch_autotune(b, 3, report, (int)sizeof report));
                    std::printf("[sweep] autotuner (%lld cells still running):\n%s", live, report);
                    tuned = live;
                }
            }
        }
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("[sweep] all cases done in %.3f s\n", wall);
        std::printf("%-5s %-28s %-11s %-10s %-14s %-14s %-14s %-9s %-14s %-14s\n", "case", "config", "grid", "steps",
                    "t", "dt", "change", "converged", "mass_flux", "thrust");
        if (!out_dir.empty()) std::filesystem::create_directories(out_dir);
        for (int i = 0; i < n; ++i) {
            mpd_solver* s = nullptr;
            Check(mpd_batch_instance(b, i, &s));
            int L = 0, M = 0, rows = 0, cols = 0;
            Check(mpd_dims(s, &L, &M, &rows, &cols));
            double mdot = 0.0, thrust = 0.0;
            Check(mpd_outlet(s, &mdot, &thrust));
            char grid[32];
            std::snprintf(grid, sizeof grid, "%dx%d", L, M);
            const std::string name = std::filesystem::path(paths[i]).filename().string();
            std::printf("%-5d %-28s %-11s %-10d %-14.6e %-14.6e %-14.6e %-9s %-14.6e %-14.6e\n", i, name.c_str(), grid,
                        res[i].total_steps, res[i].t, res[i].dt, res[i].change, res[i].converged ? "yes" : "no", mdot,
                        thrust);
            if (!out_dir.empty()) {
                char file[64];
                std::snprintf(file, sizeof file, "case_%04d.vtk", i);
                Check(mpd_write_vtk(s, (std::filesystem::path(out_dir) / file).string().c_str()));
            }
        }
        mpd_batch_destroy(b);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mpd_sweep_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
