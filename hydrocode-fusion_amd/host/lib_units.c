/*
 * lib_units.c -- the reference's per-interface solvers, on the GPU.
 *
 * Drop-ins with the reference's names and signatures for callers that use the
 * GRP solvers one interface at a time (src/include/riemann_solver.h:24-27, 60,
 * 66, 70; src/include/flux_calc.h:28).  Each call runs one thread of the batched
 * kernels of include/hydrocode_grp.h (the same device code the time step
 * uses), so it pays a device round trip per call: callers with many
 * interfaces call the hgrp_*_batch entry points instead.
 *
 * When the device call fails the error is printed on stderr and the outputs
 * are NaN (GRP_2D_flux returns -1 and leaves the fluxes as they were), so
 * that the failure cannot pass unseen.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "hydrocode_dropin.h"
#include "hydrocode_grp.h"

/* the 15-double side layout of the batch entry points */
static void side_pack(double s[15], const struct i_f_var *f)
{
    const double v[15] = {f->gamma, f->lambda_u, f->lambda_v, f->RHO, f->U, f->V, f->P,
                          f->d_rho, f->d_u, f->d_v, f->d_p, f->t_rho, f->t_u, f->t_v, f->t_p};
    memcpy(s, v, sizeof v);
}

static int failed(int rc, const char *name)
{
    if (rc)
        fprintf(stderr, "hydrogrp: %s: device call failed (%d)\n", name, rc);
    return rc != 0;
}

static void fill_nan(double *a, int n)
{
    for (int q = 0; q < n; ++q)
        a[q] = NAN;
}

/* src/riemann_solver/riemann_solver_exact_Ben.c:30-211 */
double Riemann_solver_exact(double *U_star, double *P_star, const double gammaL, const double gammaR,
                            const double u_L, const double u_R, const double p_L, const double p_R,
                            const double c_L, const double c_R, _Bool *CRW, const double eps, const double tol,
                            int N)
{
    const double in[10] = {gammaL, gammaR, u_L, u_R, p_L, p_R, c_L, c_R, eps, tol};
    double out[6];
    if (failed(hgrp_riemann_batch(in, 1, N, out), "Riemann_solver_exact"))
        fill_nan(out, 6);
    *U_star = out[0];
    *P_star = out[1];
    CRW[0] = out[2] != 0.0;
    CRW[1] = out[3] != 0.0;
    return out[4];
}

/* src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554 (single-fluid members: D, U carry rho u v p) */
void linear_GRP_solver_Edir_Q1D(double *wave_speed, double *D, double *U, double *U_star,
                                const struct i_f_var *ifv_L, const struct i_f_var *ifv_R, const double eps,
                                const double atc)
{
    double L[15], R[15], out[16];
    side_pack(L, ifv_L);
    side_pack(R, ifv_R);
    if (failed(hgrp_q1d_batch(L, R, 1, eps, atc, out), "linear_GRP_solver_Edir_Q1D"))
        fill_nan(out, 16);
    wave_speed[0] = out[0];
    wave_speed[1] = out[1];
    for (int q = 0; q < 4; ++q) {
        D[q] = out[2 + q];
        U[q] = out[6 + q];
    }
    for (int q = 0; q < 6; ++q)
        U_star[q] = out[10 + q];
}

/* src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637, the single-fluid
 * build: z_a and phi are its constants (0, slopes -0), and D, U receive all
 * six entries rho u v p z_a phi as the reference writes them (:160-166). */
void linear_GRP_solver_Edir_G2D(double *wave_speed, double *D, double *U, double *U_star,
                                const struct i_f_var *ifv_L, const struct i_f_var *ifv_R, const double eps,
                                const double atc)
{
    double L[21], R[21], out[20];
    const double zf[6] = {0.0, -0.0, -0.0, 0.0, -0.0, -0.0};
    side_pack(L, ifv_L);
    side_pack(R, ifv_R);
    memcpy(L + 15, zf, sizeof zf);
    memcpy(R + 15, zf, sizeof zf);
    if (failed(hgrp_g2d_batch(L, R, 1, eps, atc, out), "linear_GRP_solver_Edir_G2D"))
        fill_nan(out, 20);
    wave_speed[0] = out[0];
    wave_speed[1] = out[1];
    for (int q = 0; q < 6; ++q) {
        D[q] = out[2 + q];
        U[q] = out[8 + q];
        U_star[q] = out[14 + q];
    }
}

/* src/riemann_solver/linear_grp_solver_Edir.c:33-337 */
void linear_GRP_solver_Edir(double *D, double *U, const struct i_f_var *ifv_L, const struct i_f_var *ifv_R,
                            const double eps, const double atc)
{
    double L[15], R[15], out[6];
    side_pack(L, ifv_L);
    side_pack(R, ifv_R);
    if (failed(hgrp_edir_batch(L, R, 1, eps, atc, out), "linear_GRP_solver_Edir"))
        fill_nan(out, 6);
    for (int q = 0; q < 3; ++q) {
        D[q] = out[q];
        U[q] = out[3 + q];
    }
}

/* src/flux_calc/flux_solver.c:158-237.  Side effects as the reference's: the
 * grid velocity of ifv is set to 0; U, V, d_u, d_v, t_u, t_v of both sides
 * are rotated into the interface's normal frame (:162-185); the fluxes and
 * the interface values at t + tau go to ifv when the check code is 0. */
int GRP_2D_flux(struct i_f_var *ifv, struct i_f_var *ifv_R, const double tau)
{
    const double eps = config[4];
    const double n_x = ifv->n_x, n_y = ifv->n_y;
    ifv->lambda_u = 0.0;
    ifv->lambda_v = 0.0;
    double L[15], R[15], out[9];
    side_pack(L, ifv);
    side_pack(R, ifv_R);
    const int rc = hgrp_grp2d_flux_general_batch(L, R, 1, tau, eps, n_x, n_y, out);

    struct i_f_var *s[2] = {ifv, ifv_R};
    for (int k = 0; k < 2; ++k) {
        struct i_f_var *f = s[k];
        const double u = f->U * n_x + f->V * n_y, d_u = f->d_u * n_x + f->d_v * n_y;
        const double t_u = f->t_u * n_x + f->t_v * n_y;
        f->V = -f->U * n_y + f->V * n_x;
        f->d_v = -f->d_u * n_y + f->d_v * n_x;
        f->t_v = -f->t_u * n_y + f->t_v * n_x;
        f->U = u;
        f->d_u = d_u;
        f->t_u = t_u;
    }
    if (failed(rc, "GRP_2D_flux"))
        return -1;
    const int code = (int)out[0];
    if (code)
        return code;
    ifv->F_rho = out[1];
    ifv->F_u = out[2];
    ifv->F_v = out[3];
    ifv->F_e = out[4];
    ifv->RHO_int = out[5];
    ifv->U_int = out[6];
    ifv->V_int = out[7];
    ifv->P_int = out[8];
    return 0;
}
