/*
 * lf_oracle.c -- CPU oracle for the explicit Lax-Friedrichs step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this code, as the checker or as the timed
 * CPU baseline.  The product (the HIP library) never links or calls it.
 *
 * A plain-C restatement of the reference's single-rank algorithm, in the
 * reference's own layout and pass structure (six full passes per step, see
 * orc_advance).  Every function names the reference lines it follows.  All
 * arithmetic is IEEE binary64 in the reference's expression order; build with
 * -ffp-contract=off and without -ffast-math.
 */
#include "lf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oexpr.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Symbol table of a BC expression, src/bcs/per_field_bc.cpp:82-119. */
enum {
    BS_GAMMA, BS_BETA, BS_HZ0, BS_R0, BS_DZ, BS_DR, BS_R, BS_RZ, BS_Z, BS_RNB, BS_RZNB,
    BS_RHO_NB, BS_VZ_NB, BS_VR_NB, BS_VPHI_NB, BS_E_NB, BS_HZ_NB, BS_HR_NB, BS_HPHI_NB,
    BS_RHO, BS_VZ, BS_VR, BS_VPHI, BS_HZ, BS_HR, BS_HPHI, BS_E, BS_COUNT
};
static const char* const kBcNames[BS_COUNT] = {
    "gamma", "beta", "H_z0", "r_0", "dz", "dr", "r", "r_z", "z", "r_nb", "r_z_nb",
    "rho_nb", "v_z_nb", "v_r_nb", "v_phi_nb", "e_nb", "H_z_nb", "H_r_nb", "H_phi_nb",
    "rho", "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"};

/* Symbol table of an IC expression, src/ics/expression_ic.cpp:92-120. */
enum { IS_GAMMA, IS_BETA, IS_HZ0, IS_R0, IS_Z, IS_R, IS_RZ, IS_RHO, IS_VZ, IS_VR, IS_VPHI,
       IS_HZ, IS_HR, IS_HPHI, IS_E, IS_COUNT };
static const char* const kIcNames[IS_COUNT] = {"gamma", "beta", "H_z0", "r_0", "z", "r", "r_z",
                                               "rho", "v_z", "v_r", "v_phi", "H_z", "H_r",
                                               "H_phi", "e"};

typedef struct {
    int lo, hi;
    int type[8];
    double value[8];
    oexpr* ex[8];
    int has_expr, has_axis;
    /* exprtk's symbol values persist across cells, calls and steps
     * (src/bcs/per_field_bc.cpp:47-59); so do these. */
    double sym[BS_COUNT];
} oseg;

struct orc_solver {
    orc_params p;
    int L, Mmax, Mn, rows, cols;
    double dz, dy, r_0;
    double* u0[8];
    double* u[8];
    double* ph[10];
    double* prev[7];
    double *r, *r_z, *R, *dr;
    int nseg[4];
    oseg* seg[4];
};

#define AT(a, l, m) (a)[(size_t)(l) * (size_t)s->cols + (size_t)(m)]

/* ---------------------------------------------------------------- geometry */

/* short_cathode_cosine: src/geometries/short_cathode_cosine.cpp:37-71. */
static void geom_eval(const orc_solver* s, double z, double* rin, double* rout, double* drin,
                      double* drout) {
    const double* g = s->p.geom_p;
    switch (s->p.geom_type) {
        case ORC_GEOM_STRAIGHT:
            *rin = g[0];
            *rout = g[1];
            *drin = 0.0;
            *drout = 0.0;
            break;
        case ORC_GEOM_CONDIV:
            *rin = g[0] + g[1] * sin(M_PI * z);
            *rout = g[2] - g[3] * z * z;
            *drin = g[1] * M_PI * cos(M_PI * z);
            *drout = -2.0 * g[3] * z;
            break;
        default: {
            const double zc = g[0], hw = g[1], rib = g[2], ria = g[3], ro = g[4];
            const double z_start = zc - hw, z_end = zc + hw;
            const double delta_r = ria - rib;
            const double inv_width = 1.0 / (z_end - z_start);
            if (z < z_start) {
                *rin = rib;
            } else if (z < z_end) {
                const double xi = (z - z_start) * inv_width;
                *rin = rib + delta_r * 0.5 * (1.0 - cos(M_PI * xi));
            } else {
                *rin = ria;
            }
            *rout = ro;
            if (z < z_start || z >= z_end) {
                *drin = 0.0;
            } else {
                const double xi = (z - z_start) * inv_width;
                *drin = delta_r * 0.5 * M_PI * sin(M_PI * xi) * inv_width;
            }
            *drout = 0.0;
            break;
        }
    }
}

/* Grid::Grid + Grid::Build, src/grid.cpp:3-42 (single rank: l_start = m_start = 0). */
static void build_grid(orc_solver* s) {
    const int table = s->p.geom_type == ORC_GEOM_ROWS;
    double a, b, c, d;
    if (table) {
        a = s->p.geom_rows[0][1];
        b = s->p.geom_rows[1][1];
    } else {
        geom_eval(s, 0.0, &a, &b, &c, &d);
    }
    s->r_0 = (a + b) / 2.0;
    for (int l = 0; l < s->rows; ++l) {
        const int l_global = l - 1;
        const double z = l_global * s->dz;
        double rin, rout, drin, drout;
        if (table) {
            rin = s->p.geom_rows[0][l];
            rout = s->p.geom_rows[1][l];
            drin = s->p.geom_rows[2][l];
            drout = s->p.geom_rows[3][l];
        } else {
            geom_eval(s, z, &rin, &rout, &drin, &drout);
        }
        s->R[l] = rout - rin;
        s->dr[l] = s->R[l] / s->Mmax;
        for (int m = 0; m < s->cols; ++m) {
            const int m_global = m - 1;
            const double frac = m_global * s->dy;
            AT(s->r, l, m) = (1.0 - frac) * rin + frac * rout;
            AT(s->r_z, l, m) = (1.0 - frac) * drin + frac * drout;
        }
    }
}

double orc_noise_u01(unsigned long long seed, unsigned long long idx) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (idx + 1ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ExpressionIC::Apply, src/ics/expression_ic.cpp:286-324 (serial; exprtk
 * symbol values persist from cell to cell, so a forward reference reads the
 * previous cell's value and 0 at the first cell). */
static int apply_ic(orc_solver* s, char* err, int errlen) {
    static const char* defaults[8] = {"1.0", "0.0", "0.0", "0.0", "H_z0", "H_z * r_z",
                                      "(1 - 0.9 * z) * r_0 / r", "beta / (2 * (gamma - 1))"};
    const int nv = s->p.n_ic_vars;
    const int nn = IS_COUNT + nv;
    const char** names = (const char**)malloc(sizeof(char*) * (size_t)nn);
    double* vars = (double*)calloc((size_t)nn, sizeof(double));
    oexpr* ex[8] = {0};
    int ok = 1;
    for (int i = 0; i < IS_COUNT; ++i) names[i] = kIcNames[i];
    for (int i = 0; i < nv; ++i) {
        names[IS_COUNT + i] = s->p.ic_var_names[i];
        vars[IS_COUNT + i] = s->p.ic_var_values[i];
        for (int j = 0; j < IS_COUNT; ++j) {
            if (strcmp(s->p.ic_var_names[i], kIcNames[j]) == 0) {
                snprintf(err, (size_t)errlen, "IC var '%s' shadows a built-in symbol",
                         s->p.ic_var_names[i]);
                ok = 0;
            }
        }
    }
    for (int k = 0; k < 8 && ok; ++k) {
        const char* src = s->p.ic_expr[k] ? s->p.ic_expr[k] : defaults[k];
        ex[k] = oexpr_compile(src, names, nn, err, errlen);
        if (!ex[k]) ok = 0;
    }
    if (ok) {
        vars[IS_GAMMA] = s->p.gamma;
        vars[IS_BETA] = s->p.beta;
        vars[IS_HZ0] = s->p.H_z0;
        vars[IS_R0] = s->r_0;
        const double gamma = s->p.gamma;
        for (int l = 1; l < s->rows - 1; ++l) {
            const int l_global = l - 1;
            const double z = l_global * s->dz;
            for (int m = 1; m < s->cols - 1; ++m) {
                vars[IS_Z] = z;
                vars[IS_R] = AT(s->r, l, m);
                vars[IS_RZ] = AT(s->r_z, l, m);
                vars[IS_RHO] = oexpr_eval(ex[0], vars);
                if (s->p.ic_noise_amp != 0.0) {
                    const unsigned long long idx =
                        (unsigned long long)l_global * (unsigned long long)s->Mn +
                        (unsigned long long)(m - 1);
                    const double u01 = orc_noise_u01(s->p.ic_noise_seed, idx);
                    vars[IS_RHO] = vars[IS_RHO] * (1.0 + s->p.ic_noise_amp * (2.0 * u01 - 1.0));
                }
                vars[IS_VZ] = oexpr_eval(ex[1], vars);
                vars[IS_VR] = oexpr_eval(ex[2], vars);
                vars[IS_VPHI] = oexpr_eval(ex[3], vars);
                vars[IS_HZ] = oexpr_eval(ex[4], vars);
                vars[IS_HR] = oexpr_eval(ex[5], vars);
                vars[IS_HPHI] = oexpr_eval(ex[6], vars);
                vars[IS_E] = oexpr_eval(ex[7], vars);
                AT(s->ph[ORC_RHO], l, m) = vars[IS_RHO];
                AT(s->ph[ORC_VZ], l, m) = vars[IS_VZ];
                AT(s->ph[ORC_VR], l, m) = vars[IS_VR];
                AT(s->ph[ORC_VPHI], l, m) = vars[IS_VPHI];
                AT(s->ph[ORC_HZ], l, m) = vars[IS_HZ];
                AT(s->ph[ORC_HR], l, m) = vars[IS_HR];
                AT(s->ph[ORC_HPHI], l, m) = vars[IS_HPHI];
                AT(s->ph[ORC_E], l, m) = vars[IS_E];
                const double rho = vars[IS_RHO], e = vars[IS_E];
                const double hz = vars[IS_HZ], hr = vars[IS_HR], hp = vars[IS_HPHI];
                const double p = (gamma - 1.0) * rho * e;
                AT(s->ph[ORC_P], l, m) = p;
                AT(s->ph[ORC_PT], l, m) = p + 0.5 * (hz * hz + hr * hr + hp * hp);
            }
        }
    }
    for (int k = 0; k < 8; ++k) oexpr_free(ex[k]);
    free(names);
    free(vars);
    return ok;
}

/* Fields::InitConservative, src/fields.cpp:40-55. */
void orc_init_conservative(orc_solver* s) {
    double* const* ph = s->ph;
#pragma omp parallel for collapse(2)
    for (int l = 1; l < s->rows - 1; ++l) {
        for (int m = 1; m < s->cols - 1; ++m) {
            const double r = AT(s->r, l, m);
            AT(s->u0[0], l, m) = AT(ph[ORC_RHO], l, m) * r;
            AT(s->u0[1], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VZ], l, m) * r;
            AT(s->u0[2], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VR], l, m) * r;
            AT(s->u0[3], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VPHI], l, m) * r;
            AT(s->u0[4], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_E], l, m) * r;
            AT(s->u0[5], l, m) = AT(ph[ORC_HPHI], l, m);
            AT(s->u0[6], l, m) = AT(ph[ORC_HZ], l, m) * r;
            AT(s->u0[7], l, m) = AT(ph[ORC_HR], l, m) * r;
        }
    }
}

/* Fields::SavePrev, src/fields.cpp:59-74 (all cells, ghosts included). */
void orc_save_prev(orc_solver* s) {
    static const int ids[7] = {ORC_RHO, ORC_VZ, ORC_VR, ORC_VPHI, ORC_HZ, ORC_HR, ORC_HPHI};
    const size_t n = (size_t)s->rows * (size_t)s->cols;
    for (int k = 0; k < 7; ++k) memcpy(s->prev[k], s->ph[ids[k]], n * sizeof(double));
}

/* Fields::UpdatePhysicalFromU, src/fields.cpp:76-100. */
void orc_update_physical(orc_solver* s, int l_lo, int l_hi, int m_lo, int m_hi) {
    const double gamma = s->p.gamma;
    double* const* u = s->u;
    double* const* ph = s->ph;
#pragma omp parallel for collapse(2)
    for (int l = l_lo; l <= l_hi; ++l) {
        for (int m = m_lo; m <= m_hi; ++m) {
            const double r = AT(s->r, l, m);
            AT(ph[ORC_RHO], l, m) = AT(u[0], l, m) / r;
            AT(ph[ORC_VZ], l, m) = AT(u[1], l, m) / AT(u[0], l, m);
            AT(ph[ORC_VR], l, m) = AT(u[2], l, m) / AT(u[0], l, m);
            AT(ph[ORC_VPHI], l, m) = AT(u[3], l, m) / AT(u[0], l, m);
            AT(ph[ORC_HPHI], l, m) = AT(u[5], l, m);
            AT(ph[ORC_HZ], l, m) = AT(u[6], l, m) / r;
            AT(ph[ORC_HR], l, m) = AT(u[7], l, m) / r;
            AT(ph[ORC_E], l, m) = AT(u[4], l, m) / AT(u[0], l, m);
            AT(ph[ORC_P], l, m) = (gamma - 1.0) * AT(ph[ORC_RHO], l, m) * AT(ph[ORC_E], l, m);
            const double hz = AT(ph[ORC_HZ], l, m), hr = AT(ph[ORC_HR], l, m),
                         hp = AT(ph[ORC_HPHI], l, m);
            AT(ph[ORC_PT], l, m) = AT(ph[ORC_P], l, m) + 0.5 * (hz * hz + hr * hr + hp * hp);
        }
    }
}

/* Solver::ComputeCentralUpdate, src/solver.cpp:59-153 (single rank, so both
 * M boundaries are owned: m in [2, Mn-1]). */
static void central_update(orc_solver* s, double dt) {
    const int L = s->L, Mn = s->Mn;
    const double dz = s->dz;
    double* const* u0 = s->u0;
    double* const* u = s->u;
    const double* rho = s->ph[ORC_RHO];
    const double* v_z = s->ph[ORC_VZ];
    const double* v_r = s->ph[ORC_VR];
    const double* v_phi = s->ph[ORC_VPHI];
    const double* p = s->ph[ORC_P];
    const double* P = s->ph[ORC_PT];
    const double* H_z = s->ph[ORC_HZ];
    const double* H_r = s->ph[ORC_HR];
    const double* H_phi = s->ph[ORC_HPHI];
    const double* r = s->r;
    const double* dr = s->dr;
#pragma omp parallel for collapse(2)
    for (int l = 1; l <= L; ++l) {
        for (int m = 2; m <= Mn - 1; ++m) {
            const double dr_l = dr[l];
#define N(a) AT(a, l + 1, m)
#define S(a) AT(a, l - 1, m)
#define E(a) AT(a, l, m + 1)
#define W(a) AT(a, l, m - 1)
#define C(a) AT(a, l, m)
            AT(u[0], l, m) = 0.25 * (N(u0[0]) + S(u0[0]) + E(u0[0]) + W(u0[0])) +
                dt * (-(N(u0[0]) * N(v_z) - S(u0[0]) * S(v_z)) / (2 * dz) -
                      (E(u0[0]) * E(v_r) - W(u0[0]) * W(v_r)) / (2 * dr_l));

            AT(u[1], l, m) = 0.25 * (N(u0[1]) + S(u0[1]) + E(u0[1]) + W(u0[1])) +
                dt * (((N(H_z) * N(H_z) - N(P)) * N(r) - (S(H_z) * S(H_z) - S(P)) * S(r)) / (2 * dz) +
                      (E(H_z) * E(H_r) * E(r) - W(H_z) * W(H_r) * W(r)) / (2 * dr_l) -
                      (N(u0[1]) * N(v_z) - S(u0[1]) * S(v_z)) / (2 * dz) -
                      (E(u0[1]) * E(v_r) - W(u0[1]) * W(v_r)) / (2 * dr_l));

            AT(u[2], l, m) = 0.25 * (N(u0[2]) + S(u0[2]) + E(u0[2]) + W(u0[2])) +
                dt * ((C(rho) * C(v_phi) * C(v_phi) + C(P) - C(H_phi) * C(H_phi)) +
                      (N(H_z) * N(H_r) * N(r) - S(H_z) * S(H_r) * S(r)) / (2 * dz) +
                      ((E(H_r) * E(H_r) - E(P)) * E(r) - (W(H_r) * W(H_r) - W(P)) * W(r)) / (2 * dr_l) -
                      (N(u0[2]) * N(v_z) - S(u0[2]) * S(v_z)) / (2 * dz) -
                      (E(u0[2]) * E(v_r) - W(u0[2]) * W(v_r)) / (2 * dr_l));

            AT(u[3], l, m) = 0.25 * (N(u0[3]) + S(u0[3]) + E(u0[3]) + W(u0[3])) +
                dt * ((-C(rho) * C(v_r) * C(v_phi) + C(H_phi) * C(H_r)) +
                      (N(H_phi) * N(H_z) * N(r) - S(H_phi) * S(H_z) * S(r)) / (2 * dz) +
                      (E(H_phi) * E(H_r) * E(r) - W(H_phi) * W(H_r) * W(r)) / (2 * dr_l) -
                      (N(u0[3]) * N(v_z) - S(u0[3]) * S(v_z)) / (2 * dz) -
                      (E(u0[3]) * E(v_r) - W(u0[3]) * W(v_r)) / (2 * dr_l));

            AT(u[4], l, m) = 0.25 * (N(u0[4]) + S(u0[4]) + E(u0[4]) + W(u0[4])) +
                dt * (-C(p) * ((N(v_z) * N(r) - S(v_z) * S(r)) / (2 * dz) +
                               (E(v_r) * E(r) - W(v_r) * W(r)) / (2 * dr_l)) -
                      (N(u0[4]) * N(v_z) - S(u0[4]) * S(v_z)) / (2 * dz) -
                      (E(u0[4]) * E(v_r) - W(u0[4]) * W(v_r)) / (2 * dr_l));

            AT(u[5], l, m) = 0.25 * (N(u0[5]) + S(u0[5]) + E(u0[5]) + W(u0[5])) +
                dt * ((N(H_z) * N(v_phi) - S(H_z) * S(v_phi)) / (2 * dz) +
                      (E(H_r) * E(v_phi) - W(H_r) * W(v_phi)) / (2 * dr_l) -
                      (N(u0[5]) * N(v_z) - S(u0[5]) * S(v_z)) / (2 * dz) -
                      (E(u0[5]) * E(v_r) - W(u0[5]) * W(v_r)) / (2 * dr_l));

            AT(u[6], l, m) = 0.25 * (N(u0[6]) + S(u0[6]) + E(u0[6]) + W(u0[6])) +
                dt * ((E(H_r) * E(v_z) * E(r) - W(H_r) * W(v_z) * W(r)) / (2 * dr_l) -
                      (E(u0[6]) * E(v_r) - W(u0[6]) * W(v_r)) / (2 * dr_l));

            AT(u[7], l, m) = 0.25 * (N(u0[7]) + S(u0[7]) + E(u0[7]) + W(u0[7])) +
                dt * ((N(H_z) * N(v_r) * N(r) - S(H_z) * S(v_r) * S(r)) / (2 * dz) -
                      (N(u0[7]) * N(v_z) - S(u0[7]) * S(v_z)) / (2 * dz));
#undef N
#undef S
#undef E
#undef W
#undef C
        }
    }
}

/* RebuildUFromPhysical, include/physics_utils.hpp:15-24. */
static void rebuild_u(orc_solver* s, int l, int m) {
    const double r = AT(s->r, l, m);
    double* const* ph = s->ph;
    AT(s->u[0], l, m) = AT(ph[ORC_RHO], l, m) * r;
    AT(s->u[1], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VZ], l, m) * r;
    AT(s->u[2], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VR], l, m) * r;
    AT(s->u[3], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VPHI], l, m) * r;
    AT(s->u[4], l, m) = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_E], l, m) * r;
    AT(s->u[5], l, m) = AT(ph[ORC_HPHI], l, m);
    AT(s->u[6], l, m) = AT(ph[ORC_HZ], l, m) * r;
    AT(s->u[7], l, m) = AT(ph[ORC_HR], l, m) * r;
}

/* PerFieldBC::AxisLfU1/U2/U5/U7, src/bcs/per_field_bc.cpp:490-566. */
static double axis_u1(const orc_solver* s, int l, int m, double dt) {
    const double* u0 = s->u0[0];
    const double* vz = s->ph[ORC_VZ];
    const double* vr = s->ph[ORC_VR];
    const double* r = s->r;
    const double dr_l = s->dr[l], dz = s->dz;
    return (0.25 * (AT(u0, l + 1, m) / AT(r, l + 1, m) + AT(u0, l - 1, m) / AT(r, l - 1, m) +
                    AT(u0, l, m + 1) / AT(r, l, m + 1) + AT(u0, l, m) / AT(r, l, m)) +
            dt * (-(AT(u0, l + 1, m) / AT(r, l + 1, m) * AT(vz, l + 1, m) -
                    AT(u0, l - 1, m) / AT(r, l - 1, m) * AT(vz, l - 1, m)) / (2.0 * dz) -
                  (AT(u0, l, m + 1) / AT(r, l, m + 1) * AT(vr, l, m + 1) -
                   AT(u0, l, m) / AT(r, l, m) * AT(vr, l, m + 1)) / dr_l)) *
           AT(r, l, m);
}

static double axis_u2(const orc_solver* s, int l, int m, double dt) {
    const double* u0 = s->u0[1];
    const double* vz = s->ph[ORC_VZ];
    const double* vr = s->ph[ORC_VR];
    const double* Hz = s->ph[ORC_HZ];
    const double* Hr = s->ph[ORC_HR];
    const double* P = s->ph[ORC_PT];
    const double* r = s->r;
    const double dr_l = s->dr[l], dz = s->dz;
    return (0.25 * (AT(u0, l + 1, m) / AT(r, l + 1, m) + AT(u0, l - 1, m) / AT(r, l - 1, m) +
                    AT(u0, l, m + 1) / AT(r, l, m + 1) + AT(u0, l, m) / AT(r, l, m)) +
            dt * (((AT(Hz, l + 1, m) * AT(Hz, l + 1, m) - AT(P, l + 1, m)) -
                   (AT(Hz, l - 1, m) * AT(Hz, l - 1, m) - AT(P, l - 1, m))) / (2.0 * dz) +
                  (AT(Hz, l, m + 1) * AT(Hr, l, m + 1) - AT(Hz, l, m) * AT(Hr, l, m)) / dr_l -
                  (AT(u0, l + 1, m) / AT(r, l + 1, m) * AT(vz, l + 1, m) -
                   AT(u0, l - 1, m) / AT(r, l - 1, m) * AT(vz, l - 1, m)) / (2.0 * dz) -
                  (AT(u0, l, m + 1) / AT(r, l, m + 1) * AT(vr, l, m + 1) -
                   AT(u0, l, m) / AT(r, l, m) * AT(vr, l, m)) / dr_l)) *
           AT(r, l, m);
}

static double axis_u5(const orc_solver* s, int l, int m, double dt) {
    const double* u0 = s->u0[4];
    const double* vz = s->ph[ORC_VZ];
    const double* vr = s->ph[ORC_VR];
    const double* p = s->ph[ORC_P];
    const double* r = s->r;
    const double dr_l = s->dr[l], dz = s->dz;
    return (0.25 * (AT(u0, l + 1, m) / AT(r, l + 1, m) + AT(u0, l - 1, m) / AT(r, l - 1, m) +
                    AT(u0, l, m + 1) / AT(r, l, m + 1) + AT(u0, l, m) / AT(r, l, m)) +
            dt * (-AT(p, l, m) * ((AT(vz, l + 1, m) - AT(vz, l - 1, m)) / (2.0 * dz) +
                                  (AT(vr, l, m + 1) - AT(vr, l, m)) / dr_l) -
                  (AT(u0, l + 1, m) / AT(r, l + 1, m) * AT(vz, l + 1, m) -
                   AT(u0, l - 1, m) / AT(r, l - 1, m) * AT(vz, l - 1, m)) / (2.0 * dz) -
                  (AT(u0, l, m + 1) / AT(r, l, m + 1) * AT(vr, l, m + 1) -
                   AT(u0, l, m) / AT(r, l, m) * AT(vr, l, m)) / dr_l)) *
           AT(r, l, m);
}

static double axis_u7(const orc_solver* s, int l, int m, double dt) {
    const double* u0 = s->u0[6];
    const double* vz = s->ph[ORC_VZ];
    const double* vr = s->ph[ORC_VR];
    const double* Hr = s->ph[ORC_HR];
    const double* r = s->r;
    const double dr_l = s->dr[l];
    return (0.25 * (AT(u0, l + 1, m) / AT(r, l + 1, m) + AT(u0, l - 1, m) / AT(r, l - 1, m) +
                    AT(u0, l, m + 1) / AT(r, l, m + 1) + AT(u0, l, m) / AT(r, l, m)) +
            dt * ((AT(Hr, l, m + 1) * AT(vz, l, m + 1) - AT(Hr, l, m) * AT(vz, l, m)) / dr_l -
                  (AT(u0, l, m + 1) / AT(r, l, m + 1) * AT(vr, l, m + 1) -
                   AT(u0, l, m) / AT(r, l, m) * AT(vr, l, m)) / dr_l)) *
           AT(r, l, m);
}

/* Test hook: AxisLfU1/U2/U5/U7 (k = 0, 1, 4, 6: the conserved index it
 * overwrites) at cell (l, m) on the current arrays; NaN for another k. */
double orc_axis_lf(const orc_solver* s, int k, int l, int m, double dt) {
    switch (k) {
        case 0: return axis_u1(s, l, m, dt);
        case 1: return axis_u2(s, l, m, dt);
        case 4: return axis_u5(s, l, m, dt);
        case 6: return axis_u7(s, l, m, dt);
        default: return NAN;
    }
}

/* The 8 physical arrays in the BC evaluation order rho, v_z, v_r, v_phi,
 * H_z, H_r, H_phi, e (src/bcs/per_field_bc.cpp:322-360). */
static const int kBcField[8] = {ORC_RHO, ORC_VZ, ORC_VR, ORC_VPHI, ORC_HZ, ORC_HR, ORC_HPHI, ORC_E};
/* Neighbour symbols in SetNeighbors order: rho v_z v_r v_phi e H_z H_r H_phi. */
static const int kNbField[8] = {ORC_RHO, ORC_VZ, ORC_VR, ORC_VPHI, ORC_E, ORC_HZ, ORC_HR, ORC_HPHI};

/* One boundary cell of PerFieldBC::Apply, src/bcs/per_field_bc.cpp:252-484:
 * resolve the 8 physical fields in evaluation order (Neumann nb + offset,
 * Dirichlet value, Expression value, AxisLF keeps the cell value), rebuild u,
 * then apply the AxisLF overrides.  (l, m) is the boundary cell, (ln, mn) its
 * interior neighbour. */
static void bc_cell(orc_solver* s, oseg* g, int l, int m, int ln, int mn, double z_l,
                    double dt) {
    double* const* ph = s->ph;
    if (g->has_expr) {
        g->sym[BS_DR] = s->dr[l];
        g->sym[BS_R] = AT(s->r, l, m);
        g->sym[BS_RZ] = AT(s->r_z, l, m);
        g->sym[BS_Z] = z_l;
        g->sym[BS_RNB] = AT(s->r, ln, mn);
        g->sym[BS_RZNB] = AT(s->r_z, ln, mn);
        for (int k = 0; k < 8; ++k) g->sym[BS_RHO_NB + k] = AT(ph[kNbField[k]], ln, mn);
    }
    for (int k = 0; k < 8; ++k) {
        double* f = ph[kBcField[k]];
        double val;
        switch (g->type[k]) {
            case ORC_NEUMANN: val = AT(f, ln, mn) + g->value[k]; break;
            case ORC_DIRICHLET: val = g->value[k]; break;
            case ORC_EXPRESSION: val = oexpr_eval(g->ex[k], g->sym); break;
            default: val = AT(f, l, m); break;
        }
        AT(f, l, m) = val;
        if (g->has_expr) g->sym[BS_RHO + k] = val;
    }
    rebuild_u(s, l, m);
    if (g->has_axis) {
        if (g->type[0] == ORC_AXIS_LF) AT(s->u[0], l, m) = axis_u1(s, l, m, dt);
        if (g->type[1] == ORC_AXIS_LF) AT(s->u[1], l, m) = axis_u2(s, l, m, dt);
        if (g->type[7] == ORC_AXIS_LF) AT(s->u[4], l, m) = axis_u5(s, l, m, dt);
        if (g->type[4] == ORC_AXIS_LF) AT(s->u[6], l, m) = axis_u7(s, l, m, dt);
    }
}

/* FaceBC::Apply with FaceGlobalRange / ClipToLocal / the M_LO corner policy,
 * src/face_bc.cpp:56-129, single rank (the rank owns every face). */
static void apply_face(orc_solver* s, int face, double dt) {
    const int is_l = (face == 0 || face == 1);
    const int face_hi = is_l ? s->Mmax : s->L - 1;
    for (int i = 0; i < s->nseg[face]; ++i) {
        oseg* g = &s->seg[face][i];
        const int g_lo = g->lo < 0 ? 0 : g->lo;
        const int g_hi = g->hi < 0 ? face_hi : g->hi;
        const int c_lo = g_lo > 0 ? g_lo : 0;
        const int c_hi = g_hi < face_hi ? g_hi : face_hi;
        if (c_lo > c_hi) continue;
        int lo = c_lo + 1, hi = c_hi + 1;
        if (face == 2) {
            if (lo < 2) lo = 2;
            if (hi > s->L - 1) hi = s->L - 1;
            if (lo > hi) continue;
        }
        if (g->has_expr) {
            g->sym[BS_GAMMA] = s->p.gamma;
            g->sym[BS_BETA] = s->p.beta;
            g->sym[BS_HZ0] = s->p.H_z0;
            g->sym[BS_R0] = s->r_0;
            g->sym[BS_DZ] = s->dz;
            g->sym[BS_DR] = 0.0;
        }
        if (is_l) {
            const int l = (face == 0) ? 1 : s->L;
            const int ln = (face == 0) ? 2 : s->L - 1;
            const double z_l = (l - 1) * s->dz;
            for (int m = lo; m <= hi; ++m) bc_cell(s, g, l, m, ln, m, z_l, dt);
        } else {
            const int m = (face == 2) ? 1 : s->Mn;
            const int mn = (face == 2) ? 2 : s->Mn - 1;
            for (int l = lo; l <= hi; ++l) bc_cell(s, g, l, m, l, mn, (l - 1) * s->dz, dt);
        }
    }
}

/* Fields::CopyUToU0, src/fields.cpp:102-116 (OpenMP over rows, as there). */
static void copy_u_to_u0(orc_solver* s) {
    const size_t cols = (size_t)s->cols;
#pragma omp parallel for
    for (int l = 0; l < s->rows; ++l)
        for (int k = 0; k < 8; ++k) memcpy(s->u0[k] + l * cols, s->u[k] + l * cols, cols * sizeof(double));
}

/* Solver::Advance, src/solver.cpp:22-37: central update, central physical,
 * faces L_LO, M_HI, M_LO, L_HI, full physical, u0 <- u.  (The ghost exchange
 * is a no-op on one rank.) */
void orc_advance(orc_solver* s, double dt) {
    central_update(s, dt);
    orc_update_physical(s, 1, s->L, 2, s->Mn - 1);
    apply_face(s, 0, dt);
    apply_face(s, 3, dt);
    apply_face(s, 2, dt);
    apply_face(s, 1, dt);
    orc_update_physical(s, 1, s->L, 1, s->Mn);
    copy_u_to_u0(s);
}

/* Diagnostics::MaxWaveSpeed, src/diagnostics.cpp:14-43 (std::max skips NaN). */
double orc_max_wave_speed(const orc_solver* s) {
    double mx = 0.0;
    double* const* ph = s->ph;
    const double gamma = s->p.gamma;
#pragma omp parallel for collapse(2) reduction(max : mx)
    for (int l = 1; l <= s->L; ++l) {
        for (int m = 1; m <= s->Mn; ++m) {
            const double rho = AT(ph[ORC_RHO], l, m);
            const double hz = AT(ph[ORC_HZ], l, m), hr = AT(ph[ORC_HR], l, m),
                         hp = AT(ph[ORC_HPHI], l, m);
            const double vz = AT(ph[ORC_VZ], l, m), vr = AT(ph[ORC_VR], l, m);
            const double cs = sqrt(gamma * AT(ph[ORC_P], l, m) / rho);
            const double ca = sqrt((hz * hz + hr * hr + hp * hp) / rho);
            const double v = sqrt(vz * vz + vr * vr);
            const double sp = v + cs + ca;
            mx = (mx < sp) ? sp : mx;
        }
    }
    return mx;
}

/* Diagnostics::ComputeDt, src/diagnostics.cpp:49-65. */
double orc_compute_dt(const orc_solver* s, double dt_current) {
    const double speed = orc_max_wave_speed(s);
    const double dx = (s->dy < s->dz) ? s->dy : s->dz;
    const double dt_cfl = s->p.cfl * dx / (speed + 1.0e-10);
    const double dt_grown = dt_current * s->p.growth;
    const double v = (dt_grown < dt_cfl) ? dt_grown : dt_cfl;
    return (v < s->p.dt_min) ? s->p.dt_min : ((s->p.dt_max < v) ? s->p.dt_max : v);
}

/* Diagnostics::SolutionChange, src/diagnostics.cpp:71-106.  The reference
 * sums in OpenMP order; this sums rows in l order (deterministic for every
 * thread count), so it matches to rounding, not bit for bit. */
double orc_solution_change(const orc_solver* s) {
    static const int ids[7] = {ORC_RHO, ORC_VZ, ORC_VR, ORC_VPHI, ORC_HZ, ORC_HR, ORC_HPHI};
    double* rd = (double*)calloc((size_t)s->L + 2, sizeof(double));
    double* rc = (double*)calloc((size_t)s->L + 2, sizeof(double));
#pragma omp parallel for
    for (int l = 1; l <= s->L; ++l) {
        double sd = 0.0, sc = 0.0;
        for (int m = 1; m <= s->Mn; ++m) {
            for (int k = 0; k < 7; ++k) {
                const double c = AT(s->ph[ids[k]], l, m);
                const double d = c - AT(s->prev[k], l, m);
                sd += d * d;
                sc += c * c;
            }
        }
        rd[l] = sd;
        rc[l] = sc;
    }
    double g_diff = 0.0, g_curr = 0.0;
    for (int l = 1; l <= s->L; ++l) {
        g_diff += rd[l];
        g_curr += rc[l];
    }
    free(rd);
    free(rc);
    const double nd = sqrt(g_diff), nc = sqrt(g_curr);
    return (nc > 1e-15) ? (nd / nc) : nd;
}

/* The time loop of main, src/main.cpp:191-247 (probe/VTK/log output omitted:
 * they do not change the state). */
int orc_run(orc_solver* s, int max_steps, double* t_out, double* dt_out, int* converged,
            double* last_change) {
    double t = *t_out, dt = *dt_out;
    int step = 0, conv = 0;
    double change = -1.0;
    while (t < s->p.T && !conv && (max_steps <= 0 || step < max_steps)) {
        if (t + dt > s->p.T) dt = s->p.T - t;
        orc_advance(s, dt);
        t += dt;
        ++step;
        if (s->p.adaptive) dt = orc_compute_dt(s, dt);
        if (s->p.conv_threshold > 0.0 && step % s->p.check_frequency == 0) {
            change = orc_solution_change(s);
            if (change < s->p.conv_threshold) conv = 1;
            orc_save_prev(s);
        }
    }
    *t_out = t;
    *dt_out = dt;
    if (converged) *converged = conv;
    if (last_change) *last_change = change;
    return step;
}

/* Verification::ComputeIntegrals, src/verification.cpp:16-45 (serial l-major
 * order).  out: mass, mom_z, mom_r, mom_phi, energy, mag_z, mag_phi. */
void orc_integrals(const orc_solver* s, double out[7]) {
    static const int ks[7] = {0, 1, 2, 3, 4, 6, 5};
    for (int k = 0; k < 7; ++k) out[k] = 0.0;
    for (int l = 1; l <= s->L; ++l) {
        const double area = s->dr[l] * s->dz;
        for (int m = 1; m <= s->Mn; ++m) {
            for (int k = 0; k < 7; ++k) out[k] += AT(s->u0[ks[k]], l, m) * area;
        }
    }
}

/* IntegrateOutlet + GetMassFlux / GetThrust, src/integrals.cpp:44-112. */
static double outlet(const orc_solver* s, int thrust) {
    const int l = s->L;
    const double dr_out = s->dr[l];
    const double inv_8pi = 1.0 / (8.0 * M_PI);
    double sum = 0.0;
    double* const* ph = s->ph;
    for (int m = 1; m <= s->Mn; ++m) {
        const int mg = m - 1;
        const double w = (mg == 0 || mg == s->Mmax) ? 0.5 : 1.0;
        double f;
        if (thrust) {
            const double vz = AT(ph[ORC_VZ], l, m);
            const double v_z2 = vz * vz;
            const double hz = AT(ph[ORC_HZ], l, m), hr = AT(ph[ORC_HR], l, m),
                         hp = AT(ph[ORC_HPHI], l, m);
            const double H2 = hz * hz + hr * hr + hp * hp;
            f = AT(ph[ORC_RHO], l, m) * v_z2 + AT(ph[ORC_P], l, m) + H2 * inv_8pi;
        } else {
            f = AT(ph[ORC_RHO], l, m) * AT(ph[ORC_VZ], l, m);
        }
        const double F = f * 2.0 * M_PI * AT(s->r, l, m);
        sum += w * F * dr_out;
    }
    return sum;
}
double orc_mass_flux(const orc_solver* s) { return outlet(s, 0); }
double orc_thrust(const orc_solver* s) { return outlet(s, 1); }

/* Verification::CheckRadialSymmetry, src/verification.cpp:120-146. */
double orc_radial_symmetry(const orc_solver* s) {
    double mx = 0.0;
    const double* rho = s->ph[ORC_RHO];
    for (int l = 1; l <= s->L; ++l) {
        for (int m = 1; m < s->Mn; ++m) {
            const double a = AT(rho, l, m), b = AT(rho, l, m + 1);
            const double mean = 0.5 * (a + b);
            if (mean > 1.0e-15) {
                const double v = fabs(b - a) / mean;
                mx = (mx < v) ? v : mx;
            }
        }
    }
    return mx;
}

/* ------------------------------------------------------------ construction */

static void free_segs(orc_solver* s) {
    for (int f = 0; f < 4; ++f) {
        if (!s->seg[f]) continue;
        for (int i = 0; i < s->nseg[f]; ++i)
            for (int k = 0; k < 8; ++k) oexpr_free(s->seg[f][i].ex[k]);
        free(s->seg[f]);
        s->seg[f] = NULL;
    }
}

void orc_destroy(orc_solver* s) {
    if (!s) return;
    for (int k = 0; k < 8; ++k) {
        free(s->u0[k]);
        free(s->u[k]);
    }
    for (int k = 0; k < 10; ++k) free(s->ph[k]);
    for (int k = 0; k < 7; ++k) free(s->prev[k]);
    free(s->r);
    free(s->r_z);
    free(s->R);
    free(s->dr);
    free_segs(s);
    free(s);
}

orc_solver* orc_create(const orc_params* p, char* err, int errlen) {
    if (p->L_max < 3 || p->M_max < 2) {
        snprintf(err, (size_t)errlen, "grid too small");
        return NULL;
    }
    orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
    if (!s) return NULL;
    s->p = *p;
    s->L = p->L_max;
    s->Mmax = p->M_max;
    s->Mn = p->M_max + 1;
    s->rows = s->L + 2;
    s->cols = s->Mn + 2;
    /* SimConfig::Init, src/config.cpp:15-18. */
    s->dz = 1.0 / p->L_max;
    s->dy = 1.0 / p->M_max;
    const size_t n = (size_t)s->rows * (size_t)s->cols;
    int ok = 1;
    for (int k = 0; k < 8; ++k) {
        s->u0[k] = (double*)calloc(n, sizeof(double));
        s->u[k] = (double*)calloc(n, sizeof(double));
        ok = ok && s->u0[k] && s->u[k];
    }
    for (int k = 0; k < 10; ++k) ok = ok && (s->ph[k] = (double*)calloc(n, sizeof(double)));
    for (int k = 0; k < 7; ++k) ok = ok && (s->prev[k] = (double*)calloc(n, sizeof(double)));
    s->r = (double*)calloc(n, sizeof(double));
    s->r_z = (double*)calloc(n, sizeof(double));
    s->R = (double*)calloc((size_t)s->rows, sizeof(double));
    s->dr = (double*)calloc((size_t)s->rows, sizeof(double));
    ok = ok && s->r && s->r_z && s->R && s->dr;
    if (!ok) {
        snprintf(err, (size_t)errlen, "out of host memory");
        orc_destroy(s);
        return NULL;
    }
    build_grid(s);
    for (int f = 0; f < 4 && ok; ++f) {
        s->nseg[f] = p->nseg[f];
        s->seg[f] = (oseg*)calloc((size_t)(p->nseg[f] > 0 ? p->nseg[f] : 1), sizeof(oseg));
        for (int i = 0; i < p->nseg[f] && ok; ++i) {
            const orc_seg* src = &p->seg[f][i];
            oseg* g = &s->seg[f][i];
            g->lo = src->lo;
            g->hi = src->hi;
            for (int k = 0; k < 8; ++k) {
                g->type[k] = src->c[k].type;
                g->value[k] = src->c[k].value;
                if (g->type[k] == ORC_EXPRESSION) {
                    g->has_expr = 1;
                    g->ex[k] = oexpr_compile(src->c[k].expr, kBcNames, BS_COUNT, err, errlen);
                    if (!g->ex[k]) ok = 0;
                }
                if (g->type[k] == ORC_AXIS_LF) {
                    g->has_axis = 1;
                    /* PerFieldBC ctor validation, src/bcs/per_field_bc.cpp:210-234. */
                    if (f != 2) {
                        snprintf(err, (size_t)errlen, "AxisLF is only valid on the M_LO face");
                        ok = 0;
                    }
                    if (k == 2 || k == 3 || k == 5 || k == 6) {
                        snprintf(err, (size_t)errlen, "AxisLF is only valid for rho, v_z, e, H_z");
                        ok = 0;
                    }
                }
            }
        }
    }
    if (ok) ok = apply_ic(s, err, errlen);
    if (!ok) {
        orc_destroy(s);
        return NULL;
    }
    orc_init_conservative(s);
    return s;
}

int orc_rows(const orc_solver* s) { return s->rows; }
int orc_cols(const orc_solver* s) { return s->cols; }
double orc_r0(const orc_solver* s) { return s->r_0; }

double* orc_array(orc_solver* s, int kind, int k) {
    switch (kind) {
        case 0: return (k >= 0 && k < 8) ? s->u0[k] : NULL;
        case 1: return (k >= 0 && k < 8) ? s->u[k] : NULL;
        case 2: return (k >= 0 && k < 10) ? s->ph[k] : NULL;
        case 3: return (k >= 0 && k < 7) ? s->prev[k] : NULL;
        case 4: return s->r;
        case 5: return s->r_z;
        case 6: return s->dr;
        case 7: return s->R;
        default: return NULL;
    }
}
