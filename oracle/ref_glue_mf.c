/*
 * ref_glue_mf.c -- caller-side glue for the reference's interface solvers
 * built with its multi-fluid members (-DMULTIFLUID_BASICS, the switch of
 * src/include/var_struc.h:9-16) into oracle/_ref/libhydroref_mf.so.
 * TEST INFRASTRUCTURE ONLY: it produces the two-fluid interface vectors of
 * tests/golden/twofluid_vectors.npz.  It owns config[] (config[6] = gamma of
 * fluid a, config[106] = gamma of fluid b, as the reference's programs do).
 *
 * Interface sides are 21 doubles:
 *   gamma, lambda_u, lambda_v, rho, u, v, p, d_rho, d_u, d_v, d_p,
 *   t_rho, t_u, t_v, t_p, z_a, d_z_a, t_z_a, phi, d_phi, t_phi
 * (the single-fluid 15 of ref_glue.c, then the volume fraction of fluid a and
 * the mass fraction with their normal and tangential derivatives).
 */
#include <math.h>
#include <string.h>

#include "var_struc.h"
#include "riemann_solver.h"
#include "flux_calc.h"

double config[N_CONF];

void refmf_config_reset(void)
{
    for (int k = 1; k < N_CONF; ++k) config[k] = INFINITY;
    config[0] = 2.0;
}
void refmf_config_set(int idx, double val) { config[idx] = val; }

static void to_ifv(struct i_f_var *f, const double *s)
{
    memset(f, 0, sizeof *f);
    f->gamma = s[0]; f->lambda_u = s[1]; f->lambda_v = s[2];
    f->RHO = s[3]; f->U = s[4]; f->V = s[5]; f->P = s[6];
    f->d_rho = s[7]; f->d_u = s[8]; f->d_v = s[9]; f->d_p = s[10];
    f->t_rho = s[11]; f->t_u = s[12]; f->t_v = s[13]; f->t_p = s[14];
    f->Z_a = s[15]; f->d_z_a = s[16]; f->t_z_a = s[17];
    f->PHI = s[18]; f->d_phi = s[19]; f->t_phi = s[20];
}

/* linear_GRP_solver_Edir_Q1D (src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554)
 * with its multi-fluid members.  out[20]: wave_speed[2], D[6], U[6], U_star[6]
 * (D and U: rho u v p z_a phi). */
void refmf_q1d(const double *L, const double *R, double eps, double atc, double *out)
{
    struct i_f_var l, r;
    to_ifv(&l, L);
    to_ifv(&r, R);
    linear_GRP_solver_Edir_Q1D(out, out + 2, out + 8, out + 14, &l, &r, eps, atc);
}

/* linear_GRP_solver_Edir_G2D (src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637)
 * with its multi-fluid members.  out[20]: wave_speed[2], D[6], U[6], U_star[6]. */
void refmf_g2d(const double *L, const double *R, double eps, double atc, double *out)
{
    struct i_f_var l, r;
    to_ifv(&l, L);
    to_ifv(&r, R);
    linear_GRP_solver_Edir_G2D(out, out + 2, out + 8, out + 14, &l, &r, eps, atc);
}

/* GRP_2D_flux (src/flux_calc/flux_solver.c:158-237) with its multi-fluid
 * branch (:203-235): eps from config[4], the energy flux with gamma at
 * t + tau/2 from z_a there.  out[13]: return code, F_rho F_u F_v F_e, RHO_int
 * U_int V_int P_int (left side's struct after the call), Z_a and PHI at
 * t + tau, F_e_a, F_phi. */
void refmf_grp_2d_flux(const double *L, const double *R, double n_x, double n_y, double tau, double *out)
{
    struct i_f_var l, r;
    to_ifv(&l, L);
    to_ifv(&r, R);
    l.n_x = n_x; l.n_y = n_y;
    const int rc = GRP_2D_flux(&l, &r, tau);
    out[0] = rc;
    out[1] = l.F_rho; out[2] = l.F_u; out[3] = l.F_v; out[4] = l.F_e;
    out[5] = l.RHO_int; out[6] = l.U_int; out[7] = l.V_int; out[8] = l.P_int;
    out[9] = l.Z_a; out[10] = l.PHI; out[11] = l.F_e_a; out[12] = l.F_phi;
}
