// batch.hip -- batched parameter sweeps: many independent instances of the
// LF time loop (each its own config.yaml: grid size, geometry, inflow, IC)
// advanced together, one stencil launch, one boundary launch, one step-end
// launch and one convergence launch per step for all of them, K steps per
// HIP graph.
//
// Every instance keeps the semantics of a solver of its own (the reference's
// time loop, src/main.cpp:191-247, per instance: its own dt policy, T,
// step cap and convergence check); results are bit-identical to running each
// instance alone.  An instance that has finished (converged, t >= T or step
// cap) drops out of the work lists at the next graph boundary.
#include <atomic>
#include <cstdio>
#include <map>
#include <thread>
#include <tuple>

#include "solver_core.h"

namespace {

// Tile heights the autotuner tries.
const int kTileCandidates[] = {2, 3, 4, 6, 8, 12, 16, 24, 32};
// Persistent grid sizes it tries for the stencil launch (0: one workgroup per
// work entry).
const int kGridCandidates[] = {0, 512, 1024, 2048, 4096};

// Host-pinned staging buffer.
template <class T>
struct Pinned {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        HIPCHK(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

struct mpd_batch {
    std::vector<std::unique_ptr<mpd_solver>> inst;
    hipStream_t stream = nullptr;
    MpdLoopState* st = nullptr;  // one loop state per instance, contiguous
    StepArgs* d_args = nullptr;  // 2 per instance
    int* d_icur = nullptr;
    mpd::BatchAux* d_aux = nullptr;
    mpd::BatchBlock* d_work = nullptr;
    size_t work_cap = 0;
    mpd::BatchBlock* d_pwork = nullptr;  // work list of the persistent launch
    size_t pwork_cap = 0;
    Pinned<int> h_icur;
    Pinned<MpdLoopState> h_st;

    bool fast_div = true, buf = true, any_conv = false;
    std::vector<int> H;        // tile rows per instance
    int grid = 0;              // stencil workgroups (0: one per work entry)
    int steps_per_graph = 0;   // 0: chosen per run
    // 1: the steps run as one persistent launch per K steps, one workgroup
    // per instance (lf_batch_persist) instead of a graph of per-phase launches.
    int persist = 0;
    std::string tuning;        // last autotuner report

    // Work lists of the captured graph: stencil entries, boundary entries,
    // residual entries (in this order in d_work).
    std::vector<char> graph_active;
    // Steps of the graph (0..K-1) that carry the convergence launch: those on
    // which some active instance's step count reaches a check.
    std::vector<char> graph_due;
    int n_st = 0, n_bc = 0, n_res = 0;
    int bc_words = 0;  // largest BC program image of the work lists' instances
    int bc_area = 0;   // largest boundary work area (slots per lane)
    hipGraphExec_t graph = nullptr;
    int graph_K = 0;

    ~mpd_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        inst.clear();  // before the stream and the loop states they borrow
        void* bufs[] = {st, d_args, d_icur, d_aux, d_work, d_pwork};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int n() const { return (int)inst.size(); }

    void Init(const std::vector<std::string>& texts) {
        const int N = (int)texts.size();
        if (N <= 0) throw ArgError("a batch needs at least one instance");
        // Host setup (config, geometry, grid, IC, BC program) of all instances
        // in parallel: it is the slow part of construction for large sweeps.
        std::vector<std::unique_ptr<HostSetup>> hs(N);
        std::vector<std::string> errs(N);
        {
            const int nt = std::max(1, std::min(N, (int)std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            std::atomic<int> next{0};
            for (int t = 0; t < nt; ++t)
                pool.emplace_back([&] {
                    for (int i; (i = next.fetch_add(1)) < N;) {
                        try {
                            hs[i] = std::make_unique<HostSetup>(texts[i], false);
                        } catch (const std::exception& e) {
                            errs[i] = e.what();
                        }
                    }
                });
            for (auto& th : pool) th.join();
        }
        for (int i = 0; i < N; ++i)
            if (!errs[i].empty()) throw std::runtime_error("instance " + std::to_string(i) + ": " + errs[i]);
        fast_div = hs[0]->scfg.fast_div;
        for (int i = 0; i < N; ++i)
            if (hs[i]->scfg.fast_div != fast_div)
                throw std::runtime_error("instance " + std::to_string(i) +
                                         ": gpu.fast_div must be the same for every instance of a batch");

        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void**)&st, sizeof(MpdLoopState) * N));
        inst.resize(N);
        H.assign(N, 0);
        for (int i = 0; i < N; ++i) {
            inst[i] = std::make_unique<mpd_solver>();
            inst[i]->in_batch = true;
            inst[i]->InitFrom(*hs[i], stream, st + i);
            hs[i].reset();
            const mpd_solver& s = *inst[i];
            H[i] = s.scfg.tile_rows > 0 ? s.scfg.tile_rows : mpd::default_tile_rows();
            // Raw buffer addressing needs each state below 2^31 bytes.
            if (!((double)s.fstride * 8.0 * 8.0 < 2147483648.0)) buf = false;
            if (s.cfg.convergence_threshold > 0.0) any_conv = true;
        }
        HIPCHK(hipMalloc((void**)&d_args, sizeof(StepArgs) * 2 * N));
        HIPCHK(hipMalloc((void**)&d_icur, sizeof(int) * N));
        HIPCHK(hipMalloc((void**)&d_aux, sizeof(mpd::BatchAux) * N));
        h_icur.alloc(N);
        h_st.alloc(N);
        UploadArgs();
    }

    // Both ping-pong orientations of every instance; step1 is never set in a
    // batched step (an instance's first step from IC physical arrays runs on
    // its own, with the simple kernel).
    void UploadArgs() {
        std::vector<StepArgs> a(2 * n());
        std::vector<mpd::BatchAux> x(n());
        for (int i = 0; i < n(); ++i) {
            mpd_solver& s = *inst[i];
            for (int p = 0; p < 2; ++p) {
                a[2 * i + p] = s.Args(p);
                a[2 * i + p].step1 = 0;
                a[2 * i + p].phys_in = nullptr;
            }
            const bool conv = s.cfg.convergence_threshold > 0.0;
            if (conv && !s.prev) throw std::runtime_error("convergence snapshot missing");
            x[i].prev = s.prev;
            x[i].partial = s.partial;
            x[i].threshold = s.cfg.convergence_threshold;
            x[i].check_frequency = conv ? s.cfg.check_frequency : 0;
            x[i].pad = 0;
        }
        HIPCHK(hipMemcpy(d_args, a.data(), a.size() * sizeof(StepArgs), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_aux, x.data(), x.size() * sizeof(mpd::BatchAux), hipMemcpyHostToDevice));
    }

    // Residual workgroups of one instance: the single solver's partition
    // (the per-instance partial buffer holds its up to 1024 block sums).
    static int ResBlocks(const mpd_solver& s) { return mpd::residual_blocks(s.grid.L, s.grid.Mn); }

    // The batched boundary kernel uses the compact ring layout: its
    // workgroups (LDS-bound residency) then all fit at once for many small
    // instances.
    static int RingBlocks(const mpd_solver& s) { return (mpd::ring_lanes_compact(s.grid.L, s.grid.Mn) + 63) / 64; }

    // Work lists over the instances flagged in `active`.  Larger instances
    // first, so the longest chains of tiles start earliest.
    void BuildWork(const std::vector<char>& active, std::vector<mpd::BatchBlock>& out, int& nst, int& nbc,
                   int& nres) {
        std::vector<int> order;
        for (int i = 0; i < n(); ++i)
            if (active[i]) order.push_back(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return (long long)inst[a]->grid.L * inst[a]->grid.Mn > (long long)inst[b]->grid.L * inst[b]->grid.Mn;
        });
        out.clear();
        for (int i : order) {
            const mpd_solver& s = *inst[i];
            const int nstrips = mpd::stencil_strips(s.grid.Mn);
            const int ntiles = mpd::stencil_tiles(s.grid.L, s.grid.Mn, H[i]);
            for (int t = 0; t < ntiles; t += 4) out.push_back({i, t, ntiles, H[i], nstrips});
        }
        nst = (int)out.size();
        for (int i : order) {
            const int nb = RingBlocks(*inst[i]);
            for (int b = 0; b < nb; ++b) out.push_back({i, b, nb, 0, 0});
        }
        nbc = (int)out.size() - nst;
        bc_words = 0;
        bc_area = 0;
        for (int i : order) {
            bc_words = std::max(bc_words, inst[i]->bc_words);
            bc_area = std::max(bc_area, inst[i]->bc_area);
        }
        for (int i : order) {
            if (inst[i]->cfg.convergence_threshold <= 0.0) continue;
            const int nb = ResBlocks(*inst[i]);
            for (int b = 0; b < nb; ++b) out.push_back({i, b, nb, 0, 0});
        }
        nres = (int)out.size() - nst - nbc;
    }

    void UploadWork(const std::vector<mpd::BatchBlock>& w) {
        if (w.size() > work_cap) {
            if (d_work) HIPCHK(hipFree(d_work));
            d_work = nullptr;
            work_cap = std::max(w.size(), (size_t)1024);
            HIPCHK(hipMalloc((void**)&d_work, work_cap * sizeof(mpd::BatchBlock)));
        }
        if (!w.empty())
            HIPCHK(hipMemcpyAsync(d_work, w.data(), w.size() * sizeof(mpd::BatchBlock), hipMemcpyHostToDevice,
                                  stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // K steps of every active instance in one launch of the persistent
    // kernel: one work entry {instance, 0, tile count, tile rows, strips} per
    // active instance, its program image and work areas in LDS.
    void LaunchPersist(const std::vector<char>& active, int K) {
        std::vector<mpd::BatchBlock> w;
        int img = 0, area = 0;
        for (int i = 0; i < n(); ++i) {
            if (!active[i]) continue;
            const mpd_solver& s = *inst[i];
            w.push_back({i, 0, mpd::stencil_tiles(s.grid.L, s.grid.Mn, H[i]), H[i], mpd::stencil_strips(s.grid.Mn)});
            img = std::max(img, s.bc_words);
            area = std::max(area, s.bc_area);
        }
        if (w.empty()) return;
        int dev = 0, max_lds = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t lds = mpd::persist_lds_bytes(img, area);
        if (lds + 8192 > (size_t)max_lds)
            throw ArgError("persistent batch: the boundary programs need " + std::to_string(lds) +
                           " bytes of LDS per workgroup, more than the device's " + std::to_string(max_lds));
        if (w.size() > pwork_cap) {
            if (d_pwork) HIPCHK(hipFree(d_pwork));
            d_pwork = nullptr;
            pwork_cap = std::max(w.size(), (size_t)256);
            HIPCHK(hipMalloc((void**)&d_pwork, pwork_cap * sizeof(mpd::BatchBlock)));
        }
        HIPCHK(hipMemcpyAsync(d_pwork, w.data(), w.size() * sizeof(mpd::BatchBlock), hipMemcpyHostToDevice, stream));
        HIPCHK(mpd::launch_batch_persist(d_args, d_icur, d_aux, d_pwork, (int)w.size(), K, fast_div, buf, img, area,
                                         stream));
    }

    void EnqueueStep(int flip, bool check) {
        HIPCHK(mpd::launch_stencil_batch(d_args, d_icur, d_work, n_st, grid, flip, fast_div, buf, 0.0, 1, stream));
        HIPCHK(mpd::launch_bc_batch(d_args, d_icur, d_work + n_st, n_bc, flip, bc_words, bc_area, stream));
        HIPCHK(mpd::launch_finish_batch(d_args, d_icur, n(), flip, stream));
        if (check && n_res > 0)
            HIPCHK(mpd::launch_residual_batch(d_args, d_icur, d_aux, d_work + n_st + n_bc, n_res, flip, stream));
    }

    // Graph steps j in [0, K) on which an active instance checks: its step
    // count after step j is step + j + 1, and the loop checks when that is a
    // multiple of check_frequency (the residual kernel tests the same
    // condition on the device step count, so a launch left out is one that
    // would have returned at once).  An instance that stops inside the graph
    // (done, or waiting for a rescan) keeps its step count: its check, if due,
    // fell on the step where it stopped.  A rescan finished on the host moves
    // the instance's phase, which changes this set and recaptures the graph.
    std::vector<char> DueSteps(const std::vector<char>& active, int K) const {
        std::vector<char> due(K, 0);
        for (int i = 0; i < n(); ++i) {
            if (!active[i]) continue;
            const mpd_solver& s = *inst[i];
            if (s.cfg.convergence_threshold <= 0.0) continue;
            const int cf = s.cfg.check_frequency;
            if (cf <= 0) continue;
            int j = (cf - (s.step + 1) % cf) % cf;  // first j with (step + j + 1) % cf == 0
            for (; j < K; j += cf) due[j] = 1;
        }
        return due;
    }

    void Capture(const std::vector<char>& active, const std::vector<char>& due, int K) {
        if (graph) {
            HIPCHK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
        std::vector<mpd::BatchBlock> w;
        BuildWork(active, w, n_st, n_bc, n_res);
        UploadWork(w);
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int j = 0; j < K; ++j) EnqueueStep(j & 1, due[j] != 0);
        HIPCHK(hipStreamEndCapture(stream, &gr));
        const hipError_t e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
        graph_active = active;
        graph_due = due;
        graph_K = K;
    }

    int ChooseK(const std::vector<char>& active) const {
        if (steps_per_graph > 0) return std::max(2, steps_per_graph + (steps_per_graph & 1));
        // One graph per convergence check interval: an instance that converges
        // leaves the work lists at the next launch.
        int k = 0;
        for (int i = 0; i < n(); ++i) {
            if (!active[i]) continue;
            const mpd_solver& s = *inst[i];
            if (s.cfg.convergence_threshold > 0.0) k = (k == 0) ? s.cfg.check_frequency : std::min(k, s.cfg.check_frequency);
        }
        if (k == 0) k = 100;
        k = std::max(2, std::min(200, k));
        return k + (k & 1);
    }

    void ReadLoops() {
        HIPCHK(hipMemcpyAsync(h_st.p, st, sizeof(MpdLoopState) * n(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    void SwitchToIeeeDivision(int max_steps, const std::vector<int>& step0, std::vector<char>& active) {
        const int N = n();
        fast_div = false;
        for (int i = 0; i < N; ++i) {
            mpd_solver& s = *inst[i];
            if (s.scfg.fast_div) {
                s.scfg.fast_div = false;
                s.fast_div_step = s.step;
            }
            const int taken = s.step - step0[i];
            const bool capped = max_steps > 0 && taken >= max_steps;
            h_st.p[i] = s.LoopSeed(capped ? 0 : mpd_solver::Remaining(max_steps, taken));
            if (capped) h_st.p[i].done = 1;
            active[i] = (active[i] && !h_st.p[i].done) ? 1 : 0;
        }
        HIPCHK(hipMemcpyAsync(st, h_st.p, sizeof(MpdLoopState) * N, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (graph) {
            HIPCHK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
    }

    // The batched time loop: every instance runs its own loop (at most
    // max_steps more steps each, <= 0: no cap besides its time.max_steps).
    void Run(int max_steps, mpd_run_result* out) {
        const int N = n();
        // Per-instance convergence levels may have changed since the last run.
        UploadArgs();
        // An instance whose uploaded state (a warm start) is outside the split
        // division's range has switched itself to IEEE division: the batch's
        // launches cover every instance, so the batch follows.
        bool all_fast = true;
        for (int i = 0; i < N; ++i) all_fast = all_fast && inst[i]->scfg.fast_div;
        if (fast_div && !all_fast) {
            fast_div = false;
            for (int i = 0; i < N; ++i) inst[i]->scfg.fast_div = false;
            if (graph) {
                HIPCHK(hipGraphExecDestroy(graph));
                graph = nullptr;
            }
        }
        std::vector<int> step0(N);
        std::vector<char> active(N, 0);
        // Every instance's loop state in one copy.
        for (int i = 0; i < N; ++i) {
            step0[i] = inst[i]->step;
            h_st.p[i] = inst[i]->LoopSeed(max_steps);
            active[i] = h_st.p[i].done ? 0 : 1;
        }
        HIPCHK(hipMemcpyAsync(st, h_st.p, sizeof(MpdLoopState) * N, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        // First steps from IC physical arrays: each instance on its own.
        bool any1 = false;
        for (int i = 0; i < N; ++i) {
            if (!active[i] || !inst[i]->step1) continue;
            mpd_solver& s = *inst[i];
            s.EnqueueLoopStep(s.cur, s.cfg.convergence_threshold > 0.0, true);
            any1 = true;
        }
        if (any1) {
            ReadLoops();
            bool range = false;
            for (int i = 0; i < N; ++i) {
                if (!active[i] || !inst[i]->step1) continue;
                inst[i]->Absorb(h_st.p[i]);
                if (h_st.p[i].done) active[i] = 0;
                range = range || h_st.p[i].range_stop != 0;
            }
            if (range) SwitchToIeeeDivision(max_steps, step0, active);
        }
        for (;;) {
            bool any = false;
            for (int i = 0; i < N; ++i) any = any || active[i];
            if (!any) break;
            const int K = ChooseK(active);
            if (persist) {
                for (int i = 0; i < N; ++i) h_icur.p[i] = inst[i]->cur;
                HIPCHK(hipMemcpyAsync(d_icur, h_icur.p, sizeof(int) * N, hipMemcpyHostToDevice, stream));
                LaunchPersist(active, K);
            } else {
                const std::vector<char> due = DueSteps(active, K);
                if (!graph || graph_K != K || graph_active != active || graph_due != due) Capture(active, due, K);
                for (int i = 0; i < N; ++i) h_icur.p[i] = inst[i]->cur;
                HIPCHK(hipMemcpyAsync(d_icur, h_icur.p, sizeof(int) * N, hipMemcpyHostToDevice, stream));
                HIPCHK(hipGraphLaunch(graph, stream));
            }
            ReadLoops();
            bool rescans = false;
            for (int i = 0; i < N; ++i) {
                if (!active[i]) continue;
                const MpdLoopState& h = h_st.p[i];
                const int before = inst[i]->step;
                inst[i]->Absorb(h);
                if (h.done) {
                    active[i] = 0;
                } else if (h.fallback) {
                    // The step waits for its exact wave-speed rescan: finish it
                    // here; the instance rejoins the next graph.
                    inst[i]->ResolveRescan(inst[i]->cfg.convergence_threshold > 0.0);
                    rescans = true;
                } else if (inst[i]->step == before) {
                    throw std::runtime_error("batched time loop made no progress");
                }
            }
            if (rescans) {
                ReadLoops();
                for (int i = 0; i < N; ++i) {
                    if (!active[i]) continue;
                    inst[i]->Absorb(h_st.p[i]);
                    if (h_st.p[i].done) active[i] = 0;
                }
            }
            // A new state outside the split division's range (its instance
            // idles behind range_stop): the batch continues with IEEE
            // division, every loop re-seeded from its host mirror.
            bool range = false;
            for (int i = 0; i < N; ++i) range = range || h_st.p[i].range_stop != 0;
            if (range) SwitchToIeeeDivision(max_steps, step0, active);
        }
        if (out) {
            for (int i = 0; i < N; ++i) {
                const mpd_solver& s = *inst[i];
                mpd_run_result& r = out[i];
                r.t = s.t;
                r.dt = s.dt;
                r.change = s.change;
                r.steps = s.step - step0[i];
                r.total_steps = s.step;
                r.converged = s.converged ? 1 : 0;
                r.done = (!(s.t < s.cfg.T) || s.converged || (s.cfg.max_steps > 0 && s.step >= s.cfg.max_steps)) ? 1
                                                                                                               : 0;
                r.fast_div = fast_div ? 1 : 0;
            }
        }
    }

    // Average time of one batched stencil launch over all instances with the
    // current tile heights and grid (explicit dt: no loop state is read or
    // finished; the launch writes only each instance's next-state buffer and
    // row buffer, which the next step rewrites before reading).
    double TimeStencil(int reps, const std::vector<char>& live) {
        std::vector<mpd::BatchBlock> w;
        int a = 0, b = 0, c = 0;
        BuildWork(live, w, a, b, c);
        UploadWork(w);
        for (int i = 0; i < n(); ++i) h_icur.p[i] = inst[i]->cur;
        HIPCHK(hipMemcpyAsync(d_icur, h_icur.p, sizeof(int) * n(), hipMemcpyHostToDevice, stream));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const double dt = 1.0e-9;
        for (int r = 0; r < 2; ++r)
            HIPCHK(mpd::launch_stencil_batch(d_args, d_icur, d_work, a, grid, 0, fast_div, buf, dt, 0, stream));
        HIPCHK(hipEventRecord(e0, stream));
        for (int r = 0; r < reps; ++r)
            HIPCHK(mpd::launch_stencil_batch(d_args, d_icur, d_work, a, grid, 0, fast_div, buf, dt, 0, stream));
        HIPCHK(hipEventRecord(e1, stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return (double)ms / reps;
    }

    // Coordinate search: the tile height of each distinct grid size (largest
    // first, the others held), then the stencil grid, timed over the
    // instances whose loops have not ended (all of them if none is left):
    // tuned again when few instances remain, the batch's tail gets the
    // heights that suit it.  The work lists of the next run are rebuilt with
    // the result.
    void Autotune(int reps) {
        reps = std::max(1, reps);
        std::vector<char> live(n(), 0);
        bool any = false;
        for (int i = 0; i < n(); ++i) {
            live[i] = inst[i]->LoopSeed(0).done ? 0 : 1;
            any = any || live[i];
        }
        if (!any) live.assign(n(), 1);
        // Instances of one grid size share a tile height.
        std::map<std::tuple<long long, int, int>, std::vector<int>, std::greater<>> classes;
        for (int i = 0; i < n(); ++i) {
            if (!live[i]) continue;
            const mpd_solver& s = *inst[i];
            classes[{(long long)s.grid.L * s.grid.Mn, s.grid.L, s.grid.Mn}].push_back(i);
        }
        std::string rep;
        char line[256];
        double best = TimeStencil(reps, live);
        std::snprintf(line, sizeof line, "start: %.4f ms per batched stencil launch\n", best);
        rep += line;
        for (auto& kv : classes) {
            const std::vector<int>& ids = kv.second;
            const mpd_solver& s0 = *inst[ids[0]];
            int bestH = H[ids[0]];
            for (int cand : kTileCandidates) {
                if (cand > s0.grid.L) continue;
                if (cand == bestH) continue;
                for (int i : ids) H[i] = cand;
                const double t = TimeStencil(reps, live);
                if (t < best) {
                    best = t;
                    bestH = cand;
                }
            }
            for (int i : ids) H[i] = bestH;
            std::snprintf(line, sizeof line, "grid %dx%d (%zu instances): tile_rows %d -> %.4f ms\n", s0.grid.L,
                          s0.grid.Mn - 1, ids.size(), bestH, best);
            rep += line;
        }
        int bestG = grid;
        for (int g : kGridCandidates) {
            if (g == bestG) continue;
            grid = g;
            const double t = TimeStencil(reps, live);
            if (t < best) {
                best = t;
                bestG = g;
            }
        }
        grid = bestG;
        std::snprintf(line, sizeof line, "stencil grid %d (0: one workgroup per entry) -> %.4f ms\n", grid, best);
        rep += line;
        tuning = rep;
        if (graph) {
            HIPCHK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
    }
};

extern "C" {

int mpd_batch_create(const char* const* yaml_texts, int n, mpd_batch** out) {
    if (!yaml_texts || !out || n <= 0) {
        g_last_error = "mpd_batch_create: null argument or n <= 0";
        return MPD_E_ARG;
    }
    *out = nullptr;
    std::vector<std::string> texts;
    for (int i = 0; i < n; ++i) {
        if (!yaml_texts[i]) {
            g_last_error = "mpd_batch_create: null config text";
            return MPD_E_ARG;
        }
        texts.emplace_back(yaml_texts[i]);
    }
    auto b = std::make_unique<mpd_batch>();
    const int rc = Guard([&] { b->Init(texts); });
    if (rc == MPD_OK) *out = b.release();
    return rc;
}

void mpd_batch_destroy(mpd_batch* b) {
    if (b) {
        (void)hipStreamSynchronize(b->stream);
        delete b;
    }
}

int mpd_batch_size(const mpd_batch* b, int* n) {
    return Guard([&] {
        Need(b);
        Need(n);
        *n = b->n();
    });
}

int mpd_batch_instance(mpd_batch* b, int i, mpd_solver** out) {
    return Guard([&] {
        Need(b);
        Need(out);
        if (i < 0 || i >= b->n()) throw ArgError("mpd_batch_instance: index out of range");
        *out = b->inst[i].get();
    });
}

int mpd_batch_set_option(mpd_batch* b, const char* key, double value) {
    return Guard([&] {
        Need(b);
        Need(key);
        const std::string k = key;
        if (k == "steps_per_graph") {
            b->steps_per_graph = (int)value;
        } else if (k == "grid") {
            b->grid = std::max(0, (int)value);
        } else if (k == "tile_rows") {
            const int h = (int)value;
            if (h < 1) throw ArgError("tile_rows must be >= 1");
            for (int i = 0; i < b->n(); ++i) b->H[i] = h;
        } else if (k == "persistent") {
            if (value != 0.0 && value != 1.0) throw ArgError("persistent must be 0 or 1");
            b->persist = (int)value;
        } else {
            throw ArgError("unknown batch option '" + k + "'");
        }
        if (b->graph) {
            HIPCHK(hipGraphExecDestroy(b->graph));
            b->graph = nullptr;
        }
    });
}

int mpd_batch_autotune(mpd_batch* b, int reps, char* report, int report_len) {
    return Guard([&] {
        Need(b);
        b->Autotune(reps);
        if (report && report_len > 0) {
            std::snprintf(report, (size_t)report_len, "%s", b->tuning.c_str());
        }
    });
}

int mpd_batch_tile_rows(const mpd_batch* b, int i, int* tile_rows) {
    return Guard([&] {
        Need(b);
        Need(tile_rows);
        if (i < 0 || i >= b->n()) throw ArgError("mpd_batch_tile_rows: index out of range");
        *tile_rows = b->H[i];
    });
}

int mpd_batch_run(mpd_batch* b, int max_steps, mpd_run_result* results) {
    return Guard([&] {
        Need(b);
        b->Run(max_steps, results);
    });
}

}  // extern "C"
