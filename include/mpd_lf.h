/*
 * mpd_lf.h -- C ABI of the MI355X (gfx950) explicit Lax-Friedrichs solver.
 *
 * The library (mpd-plasma-dynamics_amd/libmpd_lf.so) is a drop-in for the
 * reference's time step, Solver::Advance (include/solver.hpp:24-53,
 * src/solver.cpp:22-37), and for the reductions the reference's driver runs
 * on the same Fields between steps.  The whole (z, r) domain lives on one GPU.
 *
 * Conventions: every call returns 0 on success and a negative MPD_E* code on
 * failure (mpd_last_error() gives the message of the calling thread's last
 * failure); no C++ exception crosses this boundary.  Host arrays are the
 * reference Fields layout: (L_max + 2) rows x (M_max + 3) columns of doubles,
 * row-major, ghost ring included (include/fields.hpp:10-22).  Calls on one
 * handle are stream-ordered and must come from one thread at a time; calls
 * that return numbers synchronise.
 */
#ifndef MPD_LF_H
#define MPD_LF_H

#ifdef __cplusplus
extern "C" {
#endif

#define MPD_OK 0
#define MPD_E_CONFIG -1   /* bad configuration (the reference throws std::runtime_error) */
#define MPD_E_DEVICE -2   /* HIP runtime / allocation failure (the reference aborts) */
#define MPD_E_ARG -3      /* bad argument */
#define MPD_E_STATE -4    /* non-finite or invalid state detected */

typedef struct mpd_solver mpd_solver;

/* Result of mpd_run: the loop variables of src/main.cpp:211-247. */
typedef struct {
    double t;        /* simulated time after the run */
    double dt;       /* dt variable of the loop (next step's dt before truncation) */
    double change;   /* last SolutionChange value, -1 if none was computed */
    int steps;       /* steps taken by this call */
    int total_steps; /* steps taken since creation */
    int converged;   /* SolutionChange fell below convergence.threshold */
    int done;        /* the time loop has ended: t >= T, converged, or time.max_steps reached
                        (0 when only this call's max_steps stopped it) */
    int fast_div;    /* 1: the solver still uses the split division (gpu.fast_div); 0: IEEE
                        division, by config or since a state left the split division's exact
                        range (|u_1| in [2^-30, 2^31), other conserved values 0 or in
                        [2^-100, 2^61), radii in [2^-8, 2^8]); results are IEEE-exact either way */
} mpd_run_result;

/* Build a solver from a config.yaml path / text: SimConfig::Load
 * (src/config.cpp:261-384), GeometryRegistry::Create, ExpressionIC, Grid,
 * Fields::InitPhysical/InitConservative and the Solver constructor
 * (src/main.cpp:28-185; src/solver.cpp:9-16).  The IC state is uploaded. */
int mpd_create_from_file(const char* yaml_path, mpd_solver** out);
int mpd_create_from_string(const char* yaml_text, mpd_solver** out);

/* The same with the grid's wall profile given by the caller instead of the
 * config's geometry section (which is then not read): for a geometry
 * plug-in that lives in the caller's process.  The four arrays hold, for
 * l = 0 .. L_max + 1, IGeometry's RInner, ROuter, DrInnerDz and DrOuterDz
 * (include/igeometry.hpp:16-34) at z = (l - 1) / L_max, the points
 * Grid::Build evaluates them at (src/grid.cpp:20-42); r_0 is taken from
 * row l = 1 (z = 0).  Values must be finite (MPD_E_ARG otherwise). */
int mpd_create_with_grid(const char* yaml_text, const double* r_inner, const double* r_outer,
                         const double* dr_inner_dz, const double* dr_outer_dz, mpd_solver** out);
void mpd_destroy(mpd_solver* s);
const char* mpd_last_error(void);

/* Grid sizes: L_max, M_max, rows = L_max + 2, cols = M_max + 3. */
int mpd_dims(const mpd_solver* s, int* L_max, int* M_max, int* rows, int* cols);

/* Replace the state.  u0: 8 arrays (u0_1..u0_8, Fields::u0_*); phys: 10
 * arrays in the Fields member order rho, v_z, v_r, v_phi, e, p, P, H_z, H_r,
 * H_phi, or NULL to take physical = f(u0).  When given, the next step reads
 * these physical values directly, as the reference's first step reads the IC
 * (src/main.cpp:147-148).  Ghost cells are taken as zero. */
int mpd_upload_state(mpd_solver* s, const double* const* u0, const double* const* phys);

/* Reset the time-loop variables to those of a fresh run (src/main.cpp:191-209):
 * t = 0, dt = time.dt, step count 0, no convergence check yet.  With
 * mpd_upload_state of the IC (phys given) and mpd_solution_change(s, 1, ...)
 * (the init snapshot, src/main.cpp:150-152) the next mpd_run is a cold start
 * identical to a freshly created solver's. */
int mpd_reset_loop(mpd_solver* s);

/* Copy the state out: u0 (8 arrays) and/or the physical arrays (10), either
 * may be NULL.  Physical values are those Fields holds after the last step. */
int mpd_download(mpd_solver* s, double* const* u0, double* const* phys);

/* Solver::Advance(dt), src/solver.cpp:22-37: one step with the given dt. */
int mpd_advance(mpd_solver* s, double dt);

/* Diagnostics::ComputeDt (src/diagnostics.cpp:49-65) and MaxWaveSpeed
 * (src/diagnostics.cpp:14-43) on the current state. */
int mpd_compute_dt(mpd_solver* s, double dt_current, double* dt_out);
int mpd_max_wave_speed(mpd_solver* s, double* out);

/* Diagnostics::SolutionChange (src/diagnostics.cpp:71-106) against the
 * snapshot, then Fields::SavePrev (src/fields.cpp:59-74) when save_prev. */
int mpd_solution_change(mpd_solver* s, int save_prev, double* out);

/* The time loop of src/main.cpp:211-247 (dt policy, T truncation,
 * convergence check every check_frequency steps), continued from the
 * solver's loop state, for at most max_steps steps (<= 0: no cap besides
 * time.max_steps).  Runs as device-resident HIP graphs. */
int mpd_run(mpd_solver* s, int max_steps, mpd_run_result* out);

/* Verification::ComputeIntegrals (src/verification.cpp:16-45): mass, mom_z,
 * mom_r, mom_phi, energy, mag_z, mag_phi. */
int mpd_integrals(mpd_solver* s, double out7[7]);

/* Diagnostics::GetMassFlux / GetThrust (src/integrals.cpp:44-112). */
int mpd_outlet(mpd_solver* s, double* mass_flux, double* thrust);

/* Verification::CheckRadialSymmetry (src/verification.cpp:120-146). */
int mpd_radial_symmetry(mpd_solver* s, double* out);

/* Grid arrays (Grid::Build, src/grid.cpp:20-42): r and r_z as rows x cols,
 * R and dr as rows entries, r_0.  Any pointer may be NULL. */
int mpd_grid(const mpd_solver* s, double* r, double* r_z, double* R, double* dr, double* r_0);

/* No counterpart in the reference: mpd_download for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arrays, same bits; the physical
 * arrays are staged through the library's fixed per-GPU scratch array
 * (16 MiB, held until the process ends) a chunk of rows at a time instead
 * of a buffer allocated for the call.  MPD_E_ARG when one row of the 10
 * physical fields does not fit the scratch (M_max above about 209,000).
 * Calls on one GPU share its scratch; the library serialises them. */
int mpd_download_noalloc(mpd_solver* s, double* const* u0, double* const* phys);

/* Host-only (no GPU needed): parse and validate a config (SimConfig::Load,
 * geometry, IC and BC compilation) and report the grid sizes. */
int mpd_config_check(const char* yaml_text, int* L_max, int* M_max);

/* Host-only: the step-0 state the solver starts from (ExpressionIC::Apply,
 * src/ics/expression_ic.cpp:286-324, then Fields::InitConservative,
 * src/fields.cpp:40-55) into caller arrays of rows x cols (either may be
 * NULL). */
int mpd_initial_state(const char* yaml_text, double* const* u0, double* const* phys);

/* Physical values of cell (l, m) (ghost ring included, 0 there) in the
 * Fields member order rho, v_z, v_r, v_phi, e, p, P, H_z, H_r, H_phi: what
 * the driver's checkpoint line reads (src/main.cpp:259-264). */
int mpd_probe(mpd_solver* s, int l, int m, double out10[10]);

/* IOManager::WriteFrame's VTK file (src/io_manager.cpp:218-315) of the
 * current state: legacy structured grid, binary (big-endian), points
 * (z, r, 0) over (L_max+1) x (M_max+1), arrays Rho, Energy, MPI_Rank,
 * Hphi_times_r, Velocity, MagneticField. */
int mpd_write_vtk(mpd_solver* s, const char* path);

/* Host-only: the same file from 10 host physical arrays (rows x cols, Fields
 * member order) on the grid of a config; binary != 0 for the binary form,
 * 0 for ASCII. */
int mpd_write_vtk_fields(const char* yaml_text, const double* const* phys, const char* path, int binary);

/* Kernel selection: key "kernel" (0 = simple, 1 = tiled), "fast_div" (0/1),
 * "tile_rows", "tile_order" (0: strip-fastest launch order, 1: whole row bands
 * per XCD, 2: four vertically adjacent tiles per workgroup marching in
 * alternate directions, chunk groups per XCD, 3: a group of strips per XCD
 * taken chunk by chunk), "buffers" (1: raw buffer
 * addressing for states below 2^31 bytes, the default; 0: the plain-pointer
 * path larger states take), "steps_per_graph", "convergence_threshold" (a new
 * level for the loop's convergence stop, for a config with a convergence
 * check: a run that converged continues from its snapshot to the new level;
 * a batch instance's level is taken at its batch's next run). */
int mpd_set_option(mpd_solver* s, const char* key, double value);

/* Benchmark helper: run `warmup` then `steps` loop steps with the current
 * options and return HIP-event times measured on the solver's stream:
 * total_ms for the timed steps, and per-launch averages of the stencil and
 * boundary kernels from a separately event-bracketed pass of `steps` steps. */
int mpd_time_steps(mpd_solver* s, int steps, int warmup, double* total_ms, double* stencil_ms,
                   double* bc_ms);

/* Benchmark helper, no counterpart in the reference: copies between two
 * device buffers of `mib` MiB each.  First `calib_launches` launches of one
 * fixed shape (2048 workgroups of 256, default stores; each reads and writes
 * exactly mib MiB: the calibration of the HBM traffic counters), their mean
 * time in calib_ms; then a sweep over workgroup counts and store policy
 * whose best rate (bytes read + written per second) is the copy ceiling.
 * Any output pointer may be NULL. */
int mpd_bench_copy(long long mib, int calib_launches, double* calib_ms, double* ceiling_gbs, int* best_blocks,
                   int* best_nt);

/* ------------------------------------------------------------------------
 * Batched parameter sweeps.  No counterpart in the reference, which runs one
 * case per process (src/main.cpp:22-317); a sweep over inflow parameters,
 * wall shapes and grid sizes is N such runs.  A batch holds N independent
 * instances, each built from its own config text exactly as
 * mpd_create_from_string builds one, and advances them together: one stencil
 * launch, one boundary launch and one convergence launch per step for all
 * instances (a work list of (instance, tile) entries), K steps per HIP graph.
 * Each instance's results are bit-identical to running it alone.
 */
typedef struct mpd_batch mpd_batch;

/* Build n instances from n config texts (host setup in parallel threads). */
int mpd_batch_create(const char* const* yaml_texts, int n, mpd_batch** out);
void mpd_batch_destroy(mpd_batch* b);
int mpd_batch_size(const mpd_batch* b, int* n);

/* Borrowed handle of instance i: every single-solver call above works on it
 * (upload a warm start, download, integrals, VTK ...); it lives as long as
 * the batch and mpd_destroy ignores it. */
int mpd_batch_instance(mpd_batch* b, int i, mpd_solver** out);

/* Options: "steps_per_graph" (0: the smallest convergence check_frequency of
 * the running instances; also the steps per persistent launch), "grid"
 * (stencil workgroups; 0: one per work entry, fewer: a persistent grid
 * striding over the list), "tile_rows" (all instances), "persistent" (1: each
 * run of K steps is one launch in which one workgroup per instance runs the
 * instance's steps -- central update, boundary, step end, convergence check --
 * separated by workgroup barriers; for sweeps of small instances, whose steps
 * are otherwise bound by the latency of three launches each; 0, the default:
 * a HIP graph of per-phase launches over all instances).  Results are the
 * same bits either way. */
int mpd_batch_set_option(mpd_batch* b, const char* key, double value);

/* Autotuner: times the batched stencil launch on the current states of the
 * instances whose loops have not ended (all of them if none is left; reps
 * launches per trial) and picks the tile height of each distinct grid size
 * and the stencil grid by coordinate search; writes a text report.  Called
 * again when few instances remain, it tunes the batch's tail. */
int mpd_batch_autotune(mpd_batch* b, int reps, char* report, int report_len);
int mpd_batch_tile_rows(const mpd_batch* b, int i, int* tile_rows);

/* The time loop of every instance (as mpd_run), at most max_steps more steps
 * each (<= 0: no cap besides time.max_steps), until every instance is done;
 * results: n entries, may be NULL. */
int mpd_batch_run(mpd_batch* b, int max_steps, mpd_run_result* results);

#ifdef __cplusplus
}
#endif
#endif
