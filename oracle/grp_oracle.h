/*
 * grp_oracle.h -- CPU restatement of the reference's Eulerian GRP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * Every function restates one reference routine operation for operation
 * (same expression order, no FMA contraction: build with -ffp-contract=off and
 * no -march), so that with the same libm it reproduces the reference's
 * results bit for bit.  That equality is pinned by the golden fixtures under
 * tests/golden/ (generated from the reference compiled in oracle/_ref) and by
 * closed-form checks (Toro's Sod star state, invariants) in tests/.
 *
 * Arrays are contiguous, cell (j, i) of an m x n grid at index j*n + i
 * (j = x index, i = y index), i.e. the reference's CV[k].X[j][i] with its row
 * pointers flattened.
 */
#ifndef GRP_ORACLE_H
#define GRP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_N_CONF 400

/* One side of a cell interface: the single-fluid members of the reference's
 * struct i_f_var (src/include/var_struc.h:116-142) that the GRP path reads. */
typedef struct orc_ifv {
    double gamma, lambda_u, lambda_v;
    double rho, u, v, p;
    double d_rho, d_u, d_v, d_p;   /* normal derivatives     */
    double t_rho, t_u, t_v, t_p;   /* tangential derivatives */
} orc_ifv;

/* minmod2 / minmod3: src/include/tools.h:44-73 */
static inline double orc_minmod2(const double s_L, const double s_R)
{
    if (s_L * s_R <= 0.0) return 0.0;
    if (s_R > 0.0 && s_R < s_L) return s_R;
    if (s_R <= 0.0 && s_R > s_L) return s_R;
    return s_L;
}
static inline double orc_minmod3(const double s_L, const double s_R, const double s_m)
{
    if (s_L * s_m <= 0.0 || s_R * s_m <= 0.0) return 0.0;
    if (s_m > 0.0 && s_m < s_L && s_m < s_R) return s_m;
    if (s_m <= 0.0 && s_m > s_L && s_m > s_R) return s_m;
    if (s_R > 0.0 && s_R < s_L) return s_R;
    if (s_R <= 0.0 && s_R > s_L) return s_R;
    return s_L;
}

/* Riemann_solver_exact (src/riemann_solver/riemann_solver_exact_Ben.c:30-211):
 * two-gamma exact solver, Newton on the velocity gap.  *n_iter (may be NULL)
 * receives the number of Newton iterations taken (-1 on the early exit). */
double orc_riemann_exact(double *U_star, double *P_star, double gammaL, double gammaR,
                         double u_L, double u_R, double p_L, double p_R,
                         double c_L, double c_R, int CRW[2],
                         double eps, double tol, int N, int *n_iter);

/* The same with a stopping rule: 0 the reference's (the loop above), 1 the
 * converge rule (an iterate within tol gets one more Newton step when none has
 * been taken or when its gap exceeds min(2^-49 (|v_L| + |v_R| + c_L + c_R),
 * 1e-13 p |k1 - k3|), unless an extra step gave it; no reference counterpart). */
double orc_riemann_exact_rule(double *U_star, double *P_star, double gammaL, double gammaR,
                              double u_L, double u_R, double p_L, double p_R,
                              double c_L, double c_R, int CRW[2],
                              double eps, double tol, int N, int *n_iter, int rule);

/* Newton tolerance (<= 0: eps, the reference's) and stopping rule of the
 * exact solves inside orc_grp_q1d, so of every 2D run of the oracle. */
void orc_set_newton(double tol, int rule);
extern double orc_newton_tol;
extern int orc_newton_rule;

/* Riemann_solver_exact_Ben (same file :230-408): single-gamma twin (1D path). */
double orc_riemann_exact_ben(double *U_star, double *P_star, double gamma,
                             double u_L, double u_R, double p_L, double p_R,
                             double c_L, double c_R, int CRW[2],
                             double eps, double tol, int N, int *n_iter);

/* GuessP (src/riemann_solver/riemann_solver_starPU.c:20-58): PVRS / two-shock
 * starting pressure (the named fallback guess; the 2D path never calls it). */
double orc_guessp(double DL, double DR, double UL, double UR, double PL, double PR,
                  double CL, double CR, double GammaL, double GammaR, double eps);

/* linear_GRP_solver_Edir_Q1D (src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554),
 * single-fluid members: D[0..3], U[0..3] = (rho, u, v, p); U_star[6]. */
void orc_grp_q1d(double wave_speed[2], double D[4], double U[4], double U_star[6],
                 const orc_ifv *L, const orc_ifv *R, double eps, double atc);

/* linear_GRP_solver_Edir_G2D (src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637), the
 * genuinely-2D direct Eulerian GRP solver as the reference builds it by default
 * (EXACT_TANGENT_DERIVATIVE undefined).  zf_L / zf_R: the multi-fluid members
 * z_a d_z_a t_z_a phi d_phi t_phi of each side (-DMULTIFLUID_BASICS), or NULL
 * for the single-fluid build's constants 0 -0 -0 0 -0 -0 (:70-77).  D[6], U[6] =
 * (rho, u, v, p, z_a, phi); U_star[6].  lambda_u is read from L, lambda_v from R
 * (:50).  The exact solve follows orc_set_newton like orc_grp_q1d's. */
void orc_grp_g2d(double wave_speed[2], double D[6], double U[6], double U_star[6],
                 const orc_ifv *L, const orc_ifv *R, const double *zf_L, const double *zf_R,
                 double eps, double atc);

/* ifvar_check / star_dire_check for dim == 2 (src/inter_process/fluid_var_check.c:21-51, 65-131). */
int orc_ifvar_check2(const orc_ifv *L, const orc_ifv *R, double eps);
int orc_star_dire_check2(const double mid[4], const double dire[4], double eps);

/* GRP_2D_flux (src/flux_calc/flux_solver.c:158-237).  L and R are taken by
 * value (the reference rotates them in place).  On success F[4] = (F_rho, F_u,
 * F_v, F_e) and I[4] = (RHO_int, U_int, V_int, P_int); returns 0, or the
 * star_dire_check code (1..3), in which case F and I are left untouched. */
int orc_grp_2d_flux(orc_ifv L, orc_ifv R, double tau, double eps,
                    double n_x, double n_y, double F[4], double I[4]);

/* The interface solver of orc_grp_2d_flux, so of every 2D run of the oracle:
 * 0 linear_GRP_solver_Edir_Q1D (the reference's call, flux_solver.c:191), 1
 * linear_GRP_solver_Edir_G2D with atc = eps (its commented-out alternative,
 * :189).  G2D reads lambda_v from the right side, which the flux generators
 * leave at 0. */
void orc_set_flux_solver(int solver);
extern int orc_flux_solver;

/* GRP_solver_2D_split_EUL_source (src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379).
 * rho[k], u[k], v[k], p[k], e[k] point at frame k (m*n doubles each), k < N_T;
 * frame 0 holds the initial data.  config[] has the reference's index meaning
 * (1 t_all, 4 eps, 5 N, 6 gamma, 7 CFL, 10/11 h_x/h_y, 16 tau, 17/18 BCs,
 * 41 alpha); config[5] receives the step counter on return, as in the
 * reference.  Returns 0 on a normal end, 1 on an aborting error, 2 when a
 * stop flag ended the loop early. */
int orc_split2d(int m, int n, double *const *rho, double *const *u, double *const *v,
                double *const *p, double *const *e, int N_T, int *N_plot,
                double *time_plot, double *cpu_time, double *config);

/* GRP_solver_2D_EUL_source (src/finite_volume/grp_solver_2D_EUL_source.c:75-300):
 * the unsplit step (both limiters, x and y fluxes over tau with transversal
 * derivatives, one combined update).  Arguments, config[] use and return
 * values as orc_split2d. */
int orc_unsplit2d(int m, int n, double *const *rho, double *const *u, double *const *v,
                  double *const *p, double *const *e, int N_T, int *N_plot,
                  double *time_plot, double *cpu_time, double *config);

/* linear_GRP_solver_Edir (src/riemann_solver/linear_grp_solver_Edir.c:33-337):
 * D[3], U[3] = (rho, u, p). */
void orc_grp_edir(double D[3], double U[3], const orc_ifv *L, const orc_ifv *R,
                  double eps, double atc);

/* GRP_solver_EUL_source (src/finite_volume/grp_solver_EUL_source.c:27-334).
 * rho[k] etc. point at frame k (m doubles); frame 0 = initial data. */
int orc_grp1d(int m, double *const *rho, double *const *u, double *const *p,
              double *const *e, int *N_plot, double *time_plot, double *cpu_time,
              double *config);

/* pow(x[k], y[k]) from the C library the oracle is linked with (glibc). */
void orc_pow_many(const double *x, const double *y, int n, double *out);

/* OpenMP threads the oracle's loops use (for reporting CPU baselines). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
