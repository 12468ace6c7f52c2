// Host-side setup from several threads at once, built with ThreadSanitizer
// by test_host_threads.py: a batch (mpd_batch_create) sets up its instances
// on a thread pool, and every setup registers the built-in geometries and
// creates its geometry through the process-wide registry, then parses its
// config.  Exits non-zero (TSAN's exit code) on any data race.
//
//   host_threads_check config.yaml
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <thread>
#include <vector>

#include "../mpd-plasma-dynamics_amd/host/config.h"
#include "../mpd-plasma-dynamics_amd/host/setup.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ifstream in(argv[1]);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    std::vector<std::thread> pool;
    for (int t = 0; t < 8; ++t)
        pool.emplace_back([t, &text] {
            for (int r = 0; r < 20; ++r) {
                mpd::RegisterAllGeometries();
                const mpd::YNode defaults;
                auto g = mpd::GeometryRegistry::Instance().Create(
                    (t + r) % 2 ? "straight_coaxial" : "converging_diverging", defaults);
                if (!g || mpd::GeometryRegistry::Instance().Names().size() < 3) std::abort();
                mpd::SimConfig cfg;
                cfg.LoadFromString(text);
                if (cfg.L_max <= 0) std::abort();
            }
        });
    for (auto& th : pool) th.join();
    std::puts("host_threads_check: ok");
    return 0;
}
