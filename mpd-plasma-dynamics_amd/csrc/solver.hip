// solver.hip -- the C ABI of include/mpd_lf.h: the device-resident solver that
// replaces the reference's Solver (include/solver.hpp:24-53) and the
// reductions its driver runs between steps.
#include "solver_core.h"

thread_local std::string g_last_error;

namespace {

int Create(const char* src, bool is_file, mpd_solver** out) {
    if (!src || !out) {
        g_last_error = "null argument";
        return MPD_E_ARG;
    }
    *out = nullptr;
    auto s = std::make_unique<mpd_solver>();
    const int rc = Guard([&] { s->Init(src, is_file); });
    if (rc == MPD_OK) *out = s.release();
    return rc;
}

}  // namespace

extern "C" {

int mpd_create_from_file(const char* yaml_path, mpd_solver** out) { return Create(yaml_path, true, out); }

int mpd_create_from_string(const char* yaml_text, mpd_solver** out) { return Create(yaml_text, false, out); }

int mpd_create_with_grid(const char* yaml_text, const double* r_inner, const double* r_outer,
                         const double* dr_inner_dz, const double* dr_outer_dz, mpd_solver** out) {
    if (!yaml_text || !out) {
        g_last_error = "null argument";
        return MPD_E_ARG;
    }
    *out = nullptr;
    auto s = std::make_unique<mpd_solver>();
    const HostSetup::RowGeometry rg{r_inner, r_outer, dr_inner_dz, dr_outer_dz};
    const int rc = Guard([&] { s->Init(yaml_text, false, &rg); });
    if (rc == MPD_OK) *out = s.release();
    return rc;
}

void mpd_destroy(mpd_solver* s) {
    if (s && !s->in_batch) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        delete s;
    }
}

const char* mpd_last_error(void) { return g_last_error.c_str(); }

int mpd_dims(const mpd_solver* s, int* L_max, int* M_max, int* rows, int* cols) {
    return Guard([&] {
        Need(s);
        if (L_max) *L_max = s->grid.L;
        if (M_max) *M_max = s->grid.M_max;
        if (rows) *rows = s->rows;
        if (cols) *cols = s->cols;
    });
}

int mpd_reset_loop(mpd_solver* s) {
    return Guard([&] {
        Need(s);
        s->ResetLoop();
    });
}

int mpd_upload_state(mpd_solver* s, const double* const* u0, const double* const* phys) {
    return Guard([&] {
        Need(s);
        Need(u0);
        for (int k = 0; k < 8; ++k) Need(u0[k]);
        if (phys)
            for (int k = 0; k < 10; ++k) Need(phys[k]);
        s->UploadState(u0, phys);
    });
}

int mpd_download(mpd_solver* s, double* const* u0, double* const* phys) {
    return Guard([&] {
        Need(s);
        HIPCHK(hipStreamSynchronize(s->stream));
        if (u0) s->DownloadFields(s->state[s->cur], u0, 8);
        if (phys) s->DownloadPhys(phys);
    });
}

int mpd_advance(mpd_solver* s, double dt) {
    return Guard([&] {
        Need(s);
        s->Advance(dt);
    });
}

int mpd_compute_dt(mpd_solver* s, double dt_current, double* dt_out) {
    return Guard([&] {
        Need(s);
        Need(dt_out);
        *dt_out = s->ComputeDt(dt_current);
    });
}

int mpd_max_wave_speed(mpd_solver* s, double* out) {
    return Guard([&] {
        Need(s);
        Need(out);
        *out = s->MaxWaveSpeed();
    });
}

int mpd_solution_change(mpd_solver* s, int save_prev, double* out) {
    return Guard([&] {
        Need(s);
        Need(out);
        *out = s->SolutionChange(save_prev != 0);
    });
}

int mpd_run(mpd_solver* s, int max_steps, mpd_run_result* out) {
    return Guard([&] {
        Need(s);
        s->Run(max_steps, out);
    });
}

int mpd_integrals(mpd_solver* s, double out7[7]) {
    return Guard([&] {
        Need(s);
        Need(out7);
        HIPCHK(mpd::launch_integrals(s->Args(s->cur), s->rowsum, s->scratch, s->stream));
        HIPCHK(hipMemcpyAsync(out7, s->scratch, 7 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

int mpd_outlet(mpd_solver* s, double* mass_flux, double* thrust) {
    return Guard([&] {
        Need(s);
        double v[2];
        HIPCHK(mpd::launch_outlet(s->Args(s->cur), s->grid.M_max, s->scratch, s->stream));
        HIPCHK(hipMemcpyAsync(v, s->scratch, sizeof(v), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (mass_flux) *mass_flux = v[0];
        if (thrust) *thrust = v[1];
    });
}

int mpd_radial_symmetry(mpd_solver* s, double* out) {
    return Guard([&] {
        Need(s);
        Need(out);
        unsigned long long bits = 0;
        HIPCHK(hipMemsetAsync(s->umax, 0, sizeof(bits), s->stream));
        HIPCHK(mpd::launch_radial_sym(s->Args(s->cur), s->umax, s->stream));
        HIPCHK(hipMemcpyAsync(&bits, s->umax, sizeof(bits), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        std::memcpy(out, &bits, sizeof(double));
    });
}

int mpd_grid(const mpd_solver* s, double* r, double* r_z, double* R, double* dr, double* r_0) {
    return Guard([&] {
        Need(s);
        const mpd::HostGrid& g = s->grid;
        for (int l = 0; l < g.rows; ++l) {
            for (int m = 0; m < g.cols; ++m) {
                if (r) r[(size_t)l * g.cols + m] = g.r(l, m);
                if (r_z) r_z[(size_t)l * g.cols + m] = g.r_z(l, m);
            }
            if (R) R[l] = g.R[l];
            if (dr) dr[l] = g.dr[l];
        }
        if (r_0) *r_0 = g.r_0;
    });
}

int mpd_config_check(const char* yaml_text, int* L_max, int* M_max) {
    return Guard([&] {
        Need(yaml_text);
        HostSetup h(yaml_text);
        if (L_max) *L_max = h.cfg.L_max;
        if (M_max) *M_max = h.cfg.M_max;
    });
}

int mpd_initial_state(const char* yaml_text, double* const* u0, double* const* phys) {
    return Guard([&] {
        Need(yaml_text);
        HostSetup h(yaml_text);
        const size_t n = (size_t)h.grid.rows * h.grid.cols;
        for (int k = 0; k < 8; ++k)
            if (u0 && u0[k]) std::memcpy(u0[k], h.f.u0[k].data(), n * sizeof(double));
        for (int k = 0; k < 10; ++k)
            if (phys && phys[k]) std::memcpy(phys[k], h.f.phys[k].data(), n * sizeof(double));
    });
}

int mpd_probe(mpd_solver* s, int l, int m, double out10[10]) {
    return Guard([&] {
        Need(s);
        Need(out10);
        if (l < 0 || l >= s->rows || m < 0 || m >= s->cols) throw ArgError("mpd_probe: cell outside the grid");
        HIPCHK(mpd::launch_probe(s->Args(s->cur), l, m, s->scratch, s->stream));
        HIPCHK(hipMemcpyAsync(out10, s->scratch, 10 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

int mpd_write_vtk(mpd_solver* s, const char* path) {
    return Guard([&] {
        Need(s);
        Need(path);
        HIPCHK(hipStreamSynchronize(s->stream));
        std::vector<double> ph[10];
        double* pp[10];
        for (int k = 0; k < 10; ++k) {
            ph[k].assign((size_t)s->rows * s->cols, 0.0);
            pp[k] = ph[k].data();
        }
        s->DownloadPhys(pp);
        mpd::WriteVtk(path, mpd::MakeFrame(s->cfg, s->grid, ph));
    });
}

int mpd_write_vtk_fields(const char* yaml_text, const double* const* phys, const char* path, int binary) {
    return Guard([&] {
        Need(yaml_text);
        Need(phys);
        Need(path);
        for (int k = 0; k < 10; ++k) Need(phys[k]);
        HostSetup h(yaml_text);
        const size_t n = (size_t)h.grid.rows * h.grid.cols;
        std::vector<double> ph[10];
        for (int k = 0; k < 10; ++k) ph[k].assign(phys[k], phys[k] + n);
        mpd::WriteVtk(path, mpd::MakeFrame(h.cfg, h.grid, ph), binary != 0);
    });
}

int mpd_set_option(mpd_solver* s, const char* key, double value) {
    return Guard([&] {
        Need(s);
        Need(key);
        const std::string k(key);
        if (k == "kernel") s->scfg.kind = (int)value == 0 ? mpd::STENCIL_SIMPLE : mpd::STENCIL_TILED;
        else if (k == "fast_div") s->scfg.fast_div = value != 0.0;
        else if (k == "tile_rows") s->scfg.tile_rows = (int)value;
        else if (k == "tile_order") {
            const int o = (int)value;
            if (o < 0 || o > 3) throw ArgError("tile_order must be 0, 1, 2 or 3");
            s->scfg.order = o;
        }
        else if (k == "buffers") s->scfg.buffers = value != 0.0;
        else if (k == "steps_per_graph") s->steps_per_graph = (int)value;
        else if (k == "convergence_threshold") {
            // A new stop level for the loop's convergence check: a run that
            // converged continues from its snapshot to the new level.
            if (!(value > 0.0)) throw ArgError("convergence_threshold must be positive");
            if (s->cfg.convergence_threshold <= 0.0 || !s->prev)
                throw ArgError("convergence_threshold: the config has no convergence check");
            s->cfg.convergence_threshold = value;
            s->converged = false;
        }
        else throw ArgError("unknown option '" + k + "'");
        for (auto& g : s->graph)
            if (g) {
                HIPCHK(hipGraphExecDestroy(g));
                g = nullptr;
            }
        s->graph_K = 0;
    });
}

int mpd_time_steps(mpd_solver* s, int steps, int warmup, double* total_ms, double* stencil_ms, double* bc_ms) {
    return Guard([&] {
        Need(s);
        if (steps <= 0) throw ArgError("steps must be positive");
        s->TimeSteps(steps, warmup, total_ms, stencil_ms, bc_ms);
    });
}

}  // extern "C"
