/*
 * lf_oracle.h -- CPU oracle for the explicit Lax-Friedrichs step (test
 * infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load this library, and only as the checker or the timed CPU baseline; the
 * product never links or calls it.
 *
 * It is a plain-C restatement of the reference's single-rank algorithm with
 * the reference's own data layout: separate (L+2) x (M_max+3) FP64 arrays for
 * the 8 conserved variables u0_k / u_k, the 10 physical fields and the 7
 * snapshot fields, and six full passes per step exactly as Solver::Advance
 * (src/solver.cpp:22-37).  Each function cites the reference lines it follows.
 */
#ifndef ORACLE_LF_ORACLE_H
#define ORACLE_LF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Field-condition kinds (include/field_condition.hpp:94-99). */
enum { ORC_NEUMANN = 0, ORC_DIRICHLET = 1, ORC_EXPRESSION = 2, ORC_AXIS_LF = 3 };

/* Geometry kinds: the straight coaxial channel, the converging-diverging
 * channel r_in = a0 + a1 sin(pi z), r_out = b0 - b1 z^2, the reference's
 * short_cathode_cosine (src/geometries/short_cathode_cosine.cpp:46-71), and
 * per-row wall data given by the caller (geom_rows). */
enum { ORC_GEOM_STRAIGHT = 0, ORC_GEOM_CONDIV = 1, ORC_GEOM_CATHODE_COSINE = 2, ORC_GEOM_ROWS = 3 };

/* One physical field's condition.  value: Neumann offset or Dirichlet value. */
typedef struct {
    int type;
    double value;
    const char* expr;
} orc_cond;

/* One BC segment: inclusive global range on the face's free axis (negative =
 * face start/end, src/face_bc.cpp:110-112) and eight conditions in the
 * evaluation order rho, v_z, v_r, v_phi, H_z, H_r, H_phi, e. */
typedef struct {
    int lo, hi;
    orc_cond c[8];
} orc_seg;

typedef struct {
    int L_max, M_max;
    double gamma, beta, H_z0;
    double T, dt;
    int adaptive;
    double cfl, growth, dt_min, dt_max;
    double conv_threshold;
    int check_frequency;
    int geom_type;
    double geom_p[8];
    /* IC expressions in the order rho, v_z, v_r, v_phi, H_z, H_r, H_phi, e. */
    const char* ic_expr[8];
    int n_ic_vars;
    const char* const* ic_var_names;
    const double* ic_var_values;
    /* Seeded multiplicative density noise: rho *= 1 + amp * (2U - 1). */
    double ic_noise_amp;
    unsigned long long ic_noise_seed;
    /* Faces in the order l_lo, l_hi, m_lo, m_hi. */
    int nseg[4];
    const orc_seg* seg[4];
    /* ORC_GEOM_ROWS: r_in, r_out, dr_in/dz, dr_out/dz at z = (l - 1) dz for
     * l = 0 .. L_max + 1 (the points src/grid.cpp:20-42 evaluates). */
    const double* geom_rows[4];
} orc_params;

typedef struct orc_solver orc_solver;

/* Physical field ids (the member order of include/fields.hpp:36-41). */
enum { ORC_RHO = 0, ORC_VZ, ORC_VR, ORC_VPHI, ORC_E, ORC_P, ORC_PT, ORC_HZ, ORC_HR, ORC_HPHI };

orc_solver* orc_create(const orc_params* p, char* err, int errlen);
void orc_destroy(orc_solver* s);
int orc_rows(const orc_solver* s);
int orc_cols(const orc_solver* s);

/* Array access: kind 0 = u0_k (k 0..7), 1 = u_k, 2 = physical (ids above),
 * 3 = prev (7: rho v_z v_r v_phi H_z H_r H_phi), 4 = grid r, 5 = grid r_z,
 * 6 = dr (1-D, rows entries), 7 = R (1-D). */
double* orc_array(orc_solver* s, int kind, int k);
double orc_r0(const orc_solver* s);

void orc_advance(orc_solver* s, double dt);
double orc_max_wave_speed(const orc_solver* s);
double orc_compute_dt(const orc_solver* s, double dt_current);
double orc_solution_change(const orc_solver* s);
void orc_save_prev(orc_solver* s);
void orc_update_physical(orc_solver* s, int l_lo, int l_hi, int m_lo, int m_hi);
void orc_init_conservative(orc_solver* s);

/* main.cpp's time loop (src/main.cpp:191-247) with an added step cap
 * (max_steps <= 0 means none).  Returns the number of steps taken. */
int orc_run(orc_solver* s, int max_steps, double* t_out, double* dt_out, int* converged,
            double* last_change);

void orc_integrals(const orc_solver* s, double out7[7]);
double orc_mass_flux(const orc_solver* s);
double orc_thrust(const orc_solver* s);
double orc_radial_symmetry(const orc_solver* s);

/* Test hook: PerFieldBC::AxisLfU1/U2/U5/U7 (k = 0, 1, 4, 6) at (l, m) on the
 * current u0, physical and grid arrays (src/bcs/per_field_bc.cpp:490-566). */
double orc_axis_lf(const orc_solver* s, int k, int l, int m, double dt);

/* Seeded uniform in [0,1) of the density noise: SplitMix64 of (seed, idx). */
double orc_noise_u01(unsigned long long seed, unsigned long long idx);

#ifdef __cplusplus
}
#endif
#endif
