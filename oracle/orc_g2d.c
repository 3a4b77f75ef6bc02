/*
 * orc_g2d.c -- the genuinely-2D direct Eulerian GRP solver, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_grp_g2d <- linear_GRP_solver_Edir_G2D  src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637
 *
 * The reference's default build (EXACT_TANGENT_DERIVATIVE undefined, :12-18):
 * the tangential derivatives on the t-axis come from the acoustic relations
 * across the star state (:337-349).  Both sides are handled by the same code
 * with the wave's direction sigma (-1 for the 1-wave on the left, +1 for the
 * 3-wave on the right): every mirrored expression of the reference is one of
 * x - y / x + (-y), (-a) * b / -(a * b) or -(a + b) / (-a) - b, which IEEE
 * arithmetic evaluates to the same double, signed zeros included.  Where a
 * mirror would differ in the sign of a zero (the sonic point's C), the two
 * sides are written out.
 */
#include <math.h>
#include "grp_oracle.h"

/* one side of the interface with the members the solver reads */
typedef struct {
    double sg;                           /* sigma: -1 left (1-wave), +1 right (3-wave) */
    double g, zeta, c;
    double rho, u, v, p;
    double d_rho, d_u, d_v, d_p;
    double t_rho, t_u, t_v, t_p;
    double z, d_z, t_z, phi, d_phi, t_phi;
    double rs, cs;                       /* star density and sound speed on this side */
    int crw;                             /* the wave on this side is a rarefaction */
} g2_side;

static void g2_load(g2_side *s, double sg, const orc_ifv *f, const double *zf)
{
    s->sg = sg;
    s->g = f->gamma;
    s->zeta = (f->gamma - 1.0) / (f->gamma + 1.0);
    s->rho = f->rho; s->u = f->u; s->v = f->v; s->p = f->p;
    s->d_rho = f->d_rho; s->d_u = f->d_u; s->d_v = f->d_v; s->d_p = f->d_p;
    s->t_rho = f->t_rho; s->t_u = f->t_u; s->t_v = f->t_v; s->t_p = f->t_p;
    /* without MULTIFLUID_BASICS the reference fixes these at 0 / -0 (:70-77) */
    s->z = zf ? zf[0] : 0.0;     s->d_z = zf ? zf[1] : -0.0;   s->t_z = zf ? zf[2] : -0.0;
    s->phi = zf ? zf[3] : 0.0;   s->d_phi = zf ? zf[4] : -0.0; s->t_phi = zf ? zf[5] : -0.0;
    s->c = sqrt(s->g * s->p / s->rho);
    s->crw = 0;
}

/* star density, star sound speed and the outer wave speed (:118-150) */
static double g2_star_side(g2_side *s, double p_star)
{
    if (s->crw) {
        s->rs = s->rho * pow(p_star / s->p, 1.0 / s->g);
        s->cs = s->c * pow(p_star / s->p, 0.5 * (s->g - 1.0) / s->g);
        return s->u + s->sg * s->c;
    }
    s->rs = s->rho * (p_star + s->zeta * s->p) / (s->p + s->zeta * p_star);
    s->cs = sqrt(s->g * p_star / s->rs);
    return s->u + s->sg * (s->c * sqrt(0.5 * ((s->g + 1.0) * (p_star / s->p) + (s->g - 1.0)) / s->g));
}

/* the t-axis outside the fan: upwind characteristic derivative (:159-184, :264-293) */
static void g2_upwind(double D[6], double U[6], const g2_side *s, double lu, double lv)
{
    const double a = s->u - lu, b = s->v - lv;
    U[0] = s->rho; U[1] = s->u; U[2] = s->v; U[3] = s->p; U[4] = s->z; U[5] = s->phi;
    D[0] = -a * s->d_rho - b * s->t_rho - s->rho * (s->d_u + s->t_v);
    D[1] = -a * s->d_u   - b * s->t_u   - s->d_p / s->rho;
    D[2] = -a * s->d_v   - b * s->t_v   - s->t_p / s->rho;
    D[3] = -a * s->d_p   - b * s->t_p   - s->rho * s->c * s->c * (s->d_u + s->t_v);
    D[4] = -a * s->d_z   - b * s->t_z;
    D[5] = -a * s->d_phi - b * s->t_phi;
}

/* the values at the sonic point of a rarefaction fan that contains the t-axis
 * (:188-209, :351-358, :390-397); returns C, the sound speed there */
static double g2_sonic(double U[6], const g2_side *s, double lu)
{
    double C;
    if (s->sg < 0.0) {
        U[1] = s->zeta * (s->u + 2.0 * (s->c + lu) / (s->g - 1.0));
        C = U[1] - lu;
    } else {
        U[1] = s->zeta * (s->u - 2.0 * (s->c - lu) / (s->g - 1.0));
        C = lu - U[1];
    }
    U[3] = pow(C / s->c, 2.0 * s->g / (s->g - 1.0)) * s->p;
    U[0] = s->g * U[3] / C / C;
    U[2] = s->v;
    U[4] = s->z;
    U[5] = s->phi;
    return C;
}

/* Q of the rarefaction's tangential term (:368-371): gamma = 3 has its own limit */
static double g2_q(const g2_side *s, double c_frac, double eps)
{
    const double z = s->zeta;
    if (s->g < 3.0 - eps || s->g > 3.0 + eps)
        return (c_frac * (z - 1.0) + pow(c_frac, 0.5 / z) * z) / (2.0 * z - 1.0);
    return 0.5 * c_frac + pow(c_frac, 0.5 / z) * (0.5 - 0.25 / z * log(c_frac));
}

/* the rarefaction's characteristic relation at c_frac (:362-372 and :431-441):
 * d = the right-hand side, a = 1, b = -sigma / (rho* c*) */
static double g2_crw_rhs(const g2_side *s, double c_frac, double eps)
{
    const double z = s->zeta, gm1 = s->g - 1.0;
    const double TdS = (s->d_p - s->d_rho * s->c * s->c) / gm1 / s->rho;
    const double d_R = s->d_u - s->sg * ((s->g * s->d_p / s->c - s->c * s->d_rho) / gm1 / s->rho);
    double d = ((1.0 + z) * pow(c_frac, 0.5 / z) + z * pow(c_frac, (1.0 + z) / z));
    d = d / (1.0 + 2.0 * z) * TdS;
    d = d + s->sg * (s->c * pow(c_frac, 0.5 / z) * d_R);
    return d + s->sg * (s->c * s->t_v * g2_q(s, c_frac, eps));
}

/* the t-axis inside a rarefaction fan (:351-389 left, :390-428 right) */
static void g2_in_fan(double D[6], double U[6], const g2_side *s, double lu, double lv,
                      double T_rho, double T_u, double T_p, double eps)
{
    const double C = g2_sonic(U, s, lu);
    const double z = s->zeta, gm1 = s->g - 1.0;
    const double c_frac = C / s->c;
    const double TdS = (s->d_p - s->d_rho * s->c * s->c) / gm1 / s->rho;
    const double w = U[1] - lu, b = U[2] - lv;
    D[1] = g2_crw_rhs(s, c_frac, eps);
    D[3] = U[0] * w * D[1];
    D[0] = U[0] * w * pow(c_frac, (1.0 + z) / z) * TdS * gm1;
    D[0] = (D[0] + D[3]) / C / C - b * T_rho;
    D[2] = -w * s->d_v * U[0] / s->rho - b * s->t_v - T_p / U[0];
    D[2] = D[2] - (z - 1.0) * (pow(c_frac, 2.0 / z - 1.0) - 1.0) / (z - 2.0) / U[0] * s->t_p;
    D[4] = -w * s->d_z * U[0] / s->rho - b * s->t_z;
    D[5] = -w * s->d_phi * U[0] / s->rho - b * s->t_phi;
    D[3] = D[3] - b * T_p;
    D[1] = D[1] - b * T_u + U[1] * s->t_v;
}

/* H1..H3 of a shock's characteristic relation (:445-450 / :474-479) */
static void g2_shock_h(const g2_side *s, double p_star, double H[3])
{
    const double z = s->zeta, p = s->p;
    const double VAR = sqrt((1.0 - z) / (s->rho * (p_star + z * p)));
    H[0] = 0.5 * VAR * (p_star + (1.0 + 2.0 * z) * p) / (p_star + z * p);
    H[1] = -0.5 * VAR * ((2.0 + z) * p_star + z * p) / (p_star + z * p);
    H[2] = -0.5 * VAR * (p_star - p) / s->rho;
}

/* shock speeds relative to the star state and to the unshocked side, signed
 * with the wave (:443-444 / :472-473) */
static void g2_shock_speeds(const g2_side *s, double p_star, double *SmUs, double *SmU)
{
    *SmUs = s->sg * sqrt(0.5 * ((s->g + 1.0) * s->p + (s->g - 1.0) * p_star) / s->rs);
    *SmU = s->sg * sqrt(0.5 * ((s->g + 1.0) * p_star + (s->g - 1.0) * s->p) / s->rho);
}

/* one row of the 2x2 system for (u_t, p_t) on the t-axis: a u_t + b p_t = d
 * (:421-491), with the tangential correction of its last line */
static void g2_wave_row(const g2_side *s, double p_star, double T_u, double T_p, double eps,
                        double *a, double *b, double *d)
{
    if (s->crw) {
        *a = 1.0;
        *b = -s->sg * (1.0 / s->rs / s->cs);
        *d = g2_crw_rhs(s, s->cs / s->c, eps);
    } else {
        double SmUs, SmU, H[3];
        g2_shock_speeds(s, p_star, &SmUs, &SmU);
        g2_shock_h(s, p_star, H);
        const double k = s->rho * (s->c * s->c * H[1] + H[2]);
        const double L_p = -1.0 / s->rho + s->sg * (SmU * H[1]);
        const double L_u = SmU - s->sg * k;
        const double L_rho = s->sg * (SmU * H[2]);
        const double L_v = SmUs - s->sg * k;
        *a = 1.0 + s->sg * (s->rs * SmUs * H[0]);
        *b = -(SmUs / (s->rs * s->cs * s->cs)) - s->sg * H[0];
        *d = L_rho * s->d_rho + L_u * s->d_u + L_p * s->d_p + L_v * s->t_v;
    }
    *d = *d - *a * s->v * T_u - *b * s->v * T_p;
}

void orc_grp_g2d(double wave_speed[2], double D[6], double U[6], double U_star[6],
                 const orc_ifv *Lf, const orc_ifv *Rf, const double *zf_L, const double *zf_R,
                 double eps, double atc)
{
    const double lu = Lf->lambda_u, lv = Rf->lambda_v;   /* :50 takes lambda_v from the right side */
    g2_side L, R;
    g2_load(&L, -1.0, Lf, zf_L);
    g2_load(&R, 1.0, Rf, zf_R);

    double u_star, p_star, speed_L, speed_R;
    const double dist = sqrt((L.rho - R.rho) * (L.rho - R.rho) + (L.u - R.u) * (L.u - R.u)
                             + (L.v - R.v) * (L.v - R.v) + (L.p - R.p) * (L.p - R.p));
    if (dist < atc && atc < 2 * eps) {
        u_star = 0.5 * (R.u + L.u);
        p_star = 0.5 * (R.p + L.p);
        L.rs = L.rho; L.cs = L.c; speed_L = u_star - L.cs;
        R.rs = R.rho; R.cs = R.c; speed_R = u_star + R.cs;
    } else {
        int CRW[2] = {0, 0};
        orc_riemann_exact_rule(&u_star, &p_star, L.g, R.g, L.u, R.u, L.p, R.p, L.c, R.c, CRW,
                               eps, orc_newton_tol > 0.0 ? orc_newton_tol : eps, 500, 0, orc_newton_rule);
        L.crw = CRW[0];
        R.crw = CRW[1];
        speed_L = g2_star_side(&L, p_star);
        speed_R = g2_star_side(&R, p_star);
    }
    wave_speed[0] = speed_L;
    wave_speed[1] = speed_R;
    U_star[0] = L.rs; U_star[1] = u_star; U_star[2] = R.rs;
    U_star[3] = p_star; U_star[4] = L.cs; U_star[5] = R.cs;

    if (speed_L > lu) { g2_upwind(D, U, &L, lu, lv); return; }
    if (speed_R < lu) { g2_upwind(D, U, &R, lu, lv); return; }
    const int fan_L = L.crw && ((u_star - L.cs) > lu);
    const int fan_R = R.crw && ((u_star + R.cs) < lu);
    /* the side of the contact the t-axis lies on */
    const g2_side *S = u_star > lu ? &L : &R;

    if (dist < atc) {
        /* acoustic case (:157-254) */
        double C;
        if (fan_L) C = g2_sonic(U, &L, lu);
        else if (fan_R) C = g2_sonic(U, &R, lu);
        else {
            U[0] = S->rs; U[1] = u_star; U[2] = S->v; U[3] = p_star; U[4] = S->z; U[5] = S->phi;
            C = S->cs;
        }
        const double D_p = 0.5 * ((L.d_u * (U[0] * C) + L.d_p) - (R.d_u * (U[0] * C) - R.d_p));
        const double T_p = 0.5 * ((L.t_u * (U[0] * C) + L.t_p) - (R.t_u * (U[0] * C) - R.t_p));
        const double D_u = 0.5 * (L.d_u + L.d_p / (U[0] * C) + R.d_u - R.d_p / (U[0] * C));
        const double T_u = 0.5 * (L.t_u + L.t_p / (U[0] * C) + R.t_u - R.t_p / (U[0] * C));
        const double D_rho = S->d_rho - S->d_p / (C * C) + D_p / (C * C);
        const double T_rho = S->t_rho - S->t_p / (C * C) + T_p / (C * C);
        const double w = U[1] - lu, b = U[2] - lv;
        D[0] = -w * D_rho    - b * T_rho    - U[0] * (D_u + S->t_v);
        D[1] = -w * D_u      - b * T_u      - D_p / U[0];
        D[2] = -w * S->d_v   - b * S->t_v   - T_p / U[0];
        D[3] = -w * D_p      - b * T_p      - U[0] * C * C * (D_u + S->t_v);
        D[4] = -w * S->d_z   - b * S->t_z;
        D[5] = -w * S->d_phi - b * S->t_phi;
        return;
    }

    /* tangential derivatives on the t-axis from the acoustic relations across
     * the star state of the contact's side (:337-349) */
    const g2_side *T = u_star < lu ? &R : &L;
    const double T_p = 0.5 * ((L.t_u - R.t_u) * T->rs * T->cs + L.t_p + R.t_p);
    const double T_u = 0.5 * (L.t_u + R.t_u + (L.t_p - R.t_p) / T->rs / T->cs);
    const double T_rho = T->t_rho - T->t_p / (T->cs * T->cs) + T_p / (T->cs * T->cs);

    if (fan_L) { g2_in_fan(D, U, &L, lu, lv, T_rho, T_u, T_p, eps); return; }
    if (fan_R) { g2_in_fan(D, U, &R, lu, lv, T_rho, T_u, T_p, eps); return; }

    /* non-sonic (:429-627) */
    U[0] = T->rs; U[1] = u_star; U[2] = T->v; U[3] = p_star; U[4] = T->z; U[5] = T->phi;
    const double C = T->cs;
    double a_L, b_L, d_L, a_R, b_R, d_R;
    g2_wave_row(&L, p_star, T_u, T_p, eps, &a_L, &b_L, &d_L);
    g2_wave_row(&R, p_star, T_u, T_p, eps, &a_R, &b_R, &d_R);
    const double detA = a_L * b_R - b_L * a_R;
    const double u_t_mat = (b_R * d_L - b_L * d_R) / detA;
    const double p_t_mat = (a_L * d_R - a_R * d_L) / detA;
    const double D0_p_tau = p_t_mat + U[2] * T_p;
    const double D0_u_tau = u_t_mat + U[2] * T_u;
    const double w = u_star - lu, b = U[2] - lv;
    D[1] = u_t_mat + w / U[0] / C / C * D0_p_tau + w * T->t_v;
    D[3] = p_t_mat + w * U[0] * D0_u_tau;
    if (T->crw) {
        /* the last c_frac the reference assigned: the right wave's when it is a
         * rarefaction, else the left one's (:432 and :462) */
        const double c_frac = R.crw ? R.cs / R.c : L.cs / L.c;
        const double z = T->zeta;
        D[0] = T->rs * w * pow(T->cs / T->c, (1.0 + z) / z) * (T->d_p - T->d_rho * T->c * T->c) / T->rho;
        D[0] = (D[0] + D[3] + U[2] * T_p) / T->cs / T->cs - b * T_rho;
        D[2] = -U[1] * T->d_v * U[0] / T->rho - b * T->t_v - T_p / U[0];
        D[2] = D[2] + lu * T->d_v;
        D[2] = D[2] + T->sg * (u_star / T->cs * (z - 1.0) * (pow(c_frac, 2.0 / z - 1.0) - 1.0) / (z - 2.0) / U[0]
                               * T->t_p);
        D[4] = -U[1] * T->d_z * U[0] / T->rho - b * T->t_z;
        D[4] = D[4] + lu * T->d_z;
        D[5] = -U[1] * T->d_phi * U[0] / T->rho - b * T->t_phi;
        D[5] = D[5] + lu * T->d_phi;
    } else {
        double SmUs, SmU;
        g2_shock_speeds(T, p_star, &SmUs, &SmU);
        const double z = T->zeta, r = T->rho, c = T->c;
        const double VAR = T->p + z * p_star;
        const double H1 = r * T->p * (1.0 - z * z) / VAR / VAR;
        const double H2 = r * p_star * (z * z - 1.0) / VAR / VAR;
        const double H3 = (p_star + z * T->p) / VAR;
        const double L_rho = SmU * H3 * T->d_rho;
        const double L_u = -r * (H2 * c * c + H3) * T->d_u;
        const double L_p = H2 * SmU * T->d_p;
        const double L_v = -r * (H2 * c * c + H3) * T->t_v;
        D[0] = ((u_star + SmUs) / T->cs / T->cs - u_star * H1) * D0_p_tau + T->rs * u_star * SmUs * H1 * D0_u_tau;
        D[0] = (D[0] - u_star * (L_p + L_rho + L_u + L_v)) / SmUs;
        D[0] = D[0] - b * T_rho;
        const double f = SmU * (H2 * T->d_p + H3 * T->d_rho) - r * (H2 * c * c + H3) * T->d_u;
        const double rho_x = (f + H1 * (p_t_mat - T->rs * SmUs * u_t_mat) - D[0]) / (SmU + T->u);
        D[0] = D[0] + lu * rho_x;
        D[2] = -(U[1] * (SmU * T->d_v - T->t_p / r) + (u_star + SmUs) * T_p / U[0]) / SmUs;
        D[2] = D[2] + lu * T->d_v - b * T->t_v;
        D[4] = -U[1] * SmU * T->d_z / SmUs;
        D[4] = D[4] + lu * T->d_z - b * T->t_z;
        D[5] = -U[1] * SmU * T->d_phi / SmUs;
        D[5] = D[5] + lu * T->d_phi - b * T->t_phi;
    }
    D[1] = D[1] + lv * T_u;
    D[3] = D[3] + lv * T_p;
}
