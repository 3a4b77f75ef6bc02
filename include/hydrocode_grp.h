/*
 * hydrocode_grp.h -- C ABI of the MI355X (gfx950) Eulerian GRP library.
 *
 * Plain C types only (pointers, sizes, doubles).  Two groups of entry points:
 *
 *  1. Drop-in replacements with the reference's names and signatures; see
 *     hydrocode_dropin.h (they need the reference's struct layouts).
 *
 *  2. The device-resident solver and batched interface solves that the
 *     drop-ins are built on.  They have no single counterpart in the reference:
 *     each names the reference routine whose semantics it implements.
 *
 * Every function returns 0 on success, a negative value for a HIP/runtime
 * error (-(hipError_t) or -1000 - code for argument errors), and never exits
 * the process.  All calls on one context are synchronous on return unless
 * stated otherwise; contexts are not thread-safe.
 */
#ifndef HYDROCODE_GRP_H
#define HYDROCODE_GRP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HGRP_N_CONF 400

/* ------------------------------------------------------------------------- */
/* 2D Eulerian GRP solver resident on one GPU.  Implements                    */
/* GRP_solver_2D_split_EUL_source                                             */
/* (src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379) and, when      */
/* config[33] is 0, the unsplit GRP_solver_2D_EUL_source                      */
/* (src/finite_volume/grp_solver_2D_EUL_source.c:75-300) -- the choice the    */
/* reference's main makes (src/hydrocode_2D/hydrocode.c:264-267) -- on device */
/* buffers.  Host arrays handed in/out are contiguous m*n doubles with cell   */
/* (j, i) at j*n + i, the reference's CV[k].X[j][i] with its rows packed.     */
/* ------------------------------------------------------------------------- */
typedef struct hgrp2d hgrp2d;

typedef struct hgrp2d_result {
    int    k;          /* the reference's step counter k on exit (config[5])          */
    int    nt;         /* index of the frame holding the final state (N_plot - 1)    */
    int    status;     /* 0 normal end, 1 aborted (goto return_NULL), 2 stop flag set  */
    int    sweeps;     /* directional sweeps executed                               */
    double time_c;     /* simulated time reached                                     */
    double tau;        /* last time step                                             */
    double seconds;    /* wall time of the step loop (excludes uploads)              */
} hgrp2d_result;

/* Called when the reference would write a plot frame (grp_solver_2D_split_EUL_source.c:157-176):
 * the device state is the state at that moment; the callback may call
 * hgrp2d_download().  nt: frame index the state belongs to; nt_plot: the
 * time_plot[] entry that triggered it.  Return non-zero to abort the run. */
typedef int (*hgrp2d_frame_fn)(void *user, int nt, int nt_plot, double time_c);

/* config: the reference's config[] (HGRP_N_CONF doubles, src/include/var_struc.h:44);
 * indices read: 1 t_all, 4 eps, 5 max steps, 6 gamma, 7 CFL, 10/11 h_x/h_y,
 * 16 fixed tau, 17/18 boundary conditions, 33 split (0: unsplit scheme),
 * 41 minmod alpha. */
int hgrp2d_create(hgrp2d **ctx, int m, int n, const double *config);
int hgrp2d_destroy(hgrp2d *ctx);

/* Initial data (frame 0).  e may be NULL: then E is formed as the reference's
 * main program does (src/hydrocode_2D/hydrocode.c:251-252). */
int hgrp2d_upload(hgrp2d *ctx, const double *rho, const double *u, const double *v,
                  const double *p, const double *e);
/* Current state (rho, u, v, p, E); any pointer may be NULL. */
int hgrp2d_download(hgrp2d *ctx, double *rho, double *u, double *v, double *p, double *e);

/* The reference's whole time loop: CFL tau from frame 0, X(tau/2) Y(tau)
 * X(tau/2) schedule (split) or one combined step per tau (config[33] = 0,
 * unsplit: both limiters, x- and y-interfaces with transversal derivatives, one
 * update), plot frames at time_plot[0..N_plot-1] (N_T frames kept), stop and
 * abort rules.  fn may be NULL.  For the unsplit scheme res->sweeps counts
 * steps. */
int hgrp2d_solve(hgrp2d *ctx, const double *time_plot, int N_plot, int N_T,
                 hgrp2d_frame_fn fn, void *user, hgrp2d_result *res);

/* Checkpoints (no counterpart in the reference): the device state at full
 * precision (state, both slope sets, limiter switches, the frozen h/S_max,
 * BC -1 ghosts) and, when taken inside the frame callback of hgrp2d_solve, the
 * position of the time loop.  hgrp2d_restore into a context of the same size
 * makes the next hgrp2d_solve (same time_plot) continue from that position,
 * bit for bit as if the run had not been interrupted; a checkpoint taken
 * outside a solve restarts the loop at t = 0 from the saved state.  The
 * context's own config[] is kept (t_all may be raised to run longer). */
int hgrp2d_checkpoint_size(hgrp2d *ctx, size_t *bytes);
int hgrp2d_checkpoint(hgrp2d *ctx, void *buf, size_t bytes);
int hgrp2d_restore(hgrp2d *ctx, const void *buf, size_t bytes);

/* Benchmark form: `steps` full steps X(tau/2) Y(tau) X(tau/2) with the frozen
 * tau = CFL * h/S_max, no time clipping, plot frames or host syncs inside.
 * ms_total: device time of all steps (HIP events on the solver's stream);
 * ms_x, ms_y (may be NULL): average device time of one X / one Y sweep launch,
 * from events around the launches of every sample_every-th step (NaN when
 * sample_every <= 0: then the stream holds no event but the two that time it). */
int hgrp2d_bench(hgrp2d *ctx, int steps, int sample_every, double *ms_total, double *ms_x, double *ms_y);

/* Benchmark form of the unsplit scheme: `steps` steps with tau = CFL * h/S_max.
 * ms_total: device time of all steps; ms_step (may be NULL): average device
 * time of one step launch (unsplit_step_kernel), sampled as in hgrp2d_bench. */
int hgrp2d_bench_unsplit(hgrp2d *ctx, int steps, int sample_every, double *ms_total, double *ms_step);

/* h/S_max of the uploaded initial data (the frozen CFL reduction). */
int hgrp2d_h_over_smax(hgrp2d *ctx, double *h_S_max);

/* Two gamma-law fluids (no structured counterpart in the reference; its
 * MULTIFLUID_BASICS members of GRP_2D_flux / Q1D, src/flux_calc/flux_solver.c:203-235):
 * after hgrp2d_upload, frac_a (m*n, j*n+i) gives the volume fraction of fluid a
 * (gamma = config[6]); fluid b has gamma_b (the reference's config[106]).  The
 * mixture's gamma is 1/(a/(gamma_a-1) + (1-a)/(gamma_b-1)) + 1, the volume
 * fraction is advected in quasi-conservative form, and cells with a = 1 are
 * computed exactly as in the single-fluid path.  Recomputes E from the uploaded
 * primitives and the frozen CFL bound with each cell's gamma.  Boundary conditions -2, -4, -7, -24; no
 * checkpoints in this mode. */
int hgrp2d_set_fluids(hgrp2d *ctx, double gamma_b, const double *frac_a);
int hgrp2d_download_fraction(hgrp2d *ctx, double *frac_a);

/* Stopping rules of the Newton solve for p*.
 * HGRP_NEWTON_REFERENCE: the reference's loop (riemann_solver_exact_Ben.c:133-203),
 *   which ends at the first iterate whose velocity gap is within tol.
 * HGRP_NEWTON_CONVERGE: the same loop, but an iterate whose gap is within tol
 *   gets one more Newton step if none has been taken from the starting value
 *   or if its gap is above min(2^-49 (|v_L| + |v_R| + c_L + c_R), 1e-13 p
 *   |k1 - k3|): the rounding floor of the curves' velocities, or an error above
 *   1e-13 p by the last tangents; an iterate that such an extra step gave ends
 *   the loop if its gap is within tol (else ordinary steps go on).  It ends at
 *   the root to within rounding whatever the starting guess; the reference's
 *   rule leaves up to tol / (p |dv/dp|) relative: ~1e-13 on weak waves and
 *   ~1e-10 near vacuum at tol 1e-12. */
#define HGRP_NEWTON_REFERENCE 0
#define HGRP_NEWTON_CONVERGE 1

/* Newton solve for p* (no counterpart in the reference, which always uses
 * tol = eps, its own stopping rule and its own starting guess): tol > 0 sets
 * the Newton tolerance on the velocity gap (<= 0: eps, the reference's); rule
 * is one of the HGRP_NEWTON_* rules; params (HGRP_MLP_NP floats, the master
 * vector of hgrp_mlp_train) switches on the learned starting guess, NULL
 * switches it off.  The other uses of eps are unchanged. */
int hgrp2d_set_newton(hgrp2d *ctx, double tol, int rule, const float *params);

/* The interface solver of GRP_2D_flux in the unsplit step (config[33] = 0):
 * HGRP_SOLVER_Q1D, linear_GRP_solver_Edir_Q1D, the reference's call
 * (src/flux_calc/flux_solver.c:191, the default), or HGRP_SOLVER_G2D,
 * linear_GRP_solver_Edir_G2D at atc = eps, the reference's commented-out
 * alternative on the line above it (:189; the genuinely-2D solver, see
 * hgrp_g2d_batch).  A split context (config[33] != 0) takes Q1D only: -1002
 * and a message for G2D.  Works with every Newton setting of hgrp2d_set_newton. */
/* Launch form of the unsplit step (no counterpart in the reference): TILE
 * blocks of 12 waves over 10 x 62 updated cells, STRIP one wave per 62-cell
 * strip walking down its rows; AUTO takes the tile, the faster at every size
 * measured on MI355X.  hgrp2d_tune on an unsplit context times both and keeps
 * the faster.  Results do not depend on the form. */
#define HGRP_UNSPLIT_AUTO 0
#define HGRP_UNSPLIT_TILE 1
#define HGRP_UNSPLIT_STRIP 2
int hgrp2d_set_unsplit_form(hgrp2d *ctx, int form);

#define HGRP_SOLVER_Q1D 0
#define HGRP_SOLVER_G2D 1
int hgrp2d_set_interface_solver(hgrp2d *ctx, int solver);

/* Launch variants of the X sweep (0..4) and the Y sweep (0..4) (no counterpart
 * in the reference).  X: 0 = 8 cell rows per block at 4 waves/SIMD (the
 * default), 1 = 8 rows at the compiler's register budget, 2 = 16 rows, 3 and 4
 * = 8 rows walking 4 and 8 chunks along the lines (fewer repeated solves).  Y: 0 =
 * 4 rows of 16 lanes per wave (the default), 1 = 2 x 32, 2 = 1 x 64, 3 = 8 x 8,
 * 4 = 1 x 64 at the compiler's register budget.  Results do not depend on them. */
int hgrp2d_set_sweeps(hgrp2d *ctx, int x_variant, int y_variant);

/* Sweep autotuner: times every X variant, then every Y variant, over `steps`
 * full steps from the context's current state (two trials each, per-launch
 * device time), sets the fastest with hgrp2d_set_sweeps and restores the state,
 * so that the run goes on exactly as without it.  x_variant, y_variant, ms[2]
 * (average ms per X and per Y launch of the winners) may be NULL.  Not for
 * two-fluid contexts.  On an unsplit context (config[33] = 0) it times the
 * unsplit step's two launch forms instead and sets the faster with
 * hgrp2d_set_unsplit_form: x_variant receives the form (HGRP_UNSPLIT_TILE or
 * _STRIP), y_variant -1, ms[0] the winner's average ms per step, ms[1] NaN. */
int hgrp2d_tune(hgrp2d *ctx, int steps, int *x_variant, int *y_variant, double *ms);

/* Name of the kernel the context launches for an x (dir 0) or y (dir 1) sweep
 * (it depends on the launch variant, the Newton mode and, with two fluids,
 * names the two-fluid sweeps), or for an unsplit step (dir 2); "" for a bad
 * argument.  For profiling tools. */
const char *hgrp2d_sweep_kernel(hgrp2d *ctx, int dir);

/* ------------------------------------------------------------------------- */
/* A batch of independent 2D split problems of uneven sizes, advanced        */
/* together: each sweep of every grid of the batch is one launch (no         */
/* counterpart in the reference).  Every grid runs                           */
/* GRP_solver_2D_split_EUL_source (src/finite_volume/                        */
/* grp_solver_2D_split_EUL_source.c:75-379) with plot times {0, t_all} and   */
/* its own frozen tau = CFL h/S_max of its initial data: the same steps,     */
/* state, step counter, time and status as hgrp2d_solve on that grid alone.  */
/* ------------------------------------------------------------------------- */
typedef struct hgrp2db hgrp2db;

/* count grids; grid g has m[g] x n[g] cells of size h_x[g] x h_y[g].  config:
 * the reference's config[] shared by all grids (indices read as for
 * hgrp2d_create; 10 and 11 are taken from h_x, h_y; 17 and 18 must be valid
 * boundary conditions, else ARG_ERROR - 4; t_all (1) or the step cap (5) must
 * be finite, else ARG_ERROR - 3). */
int hgrp2db_create(hgrp2db **ctx, int count, const int *m, const int *n, const double *h_x,
                   const double *h_y, const double *config);
int hgrp2db_destroy(hgrp2db *ctx);

/* Initial data: the grids' m[g] x n[g] arrays (row j, column i at j*n + i)
 * one after the other.  Computes each grid's h/S_max and its schedule. */
int hgrp2db_upload(hgrp2db *ctx, const double *rho, const double *u, const double *v, const double *p);

/* As hgrp2d_set_newton, for every grid. */
int hgrp2db_set_newton(hgrp2db *ctx, double tol, int rule, const float *params);

/* Launch variants as hgrp2d_set_sweeps (the same numbering). */
int hgrp2db_set_sweeps(hgrp2db *ctx, int x_variant, int y_variant);

/* Times every X variant, then every Y variant, over the first `steps` steps
 * of the batch from the uploaded state (two trials each), keeps the fastest
 * and restores the state.  ms[2]: average ms per X and per Y launch of the
 * winners (any of the outputs may be NULL). */
int hgrp2db_tune(hgrp2db *ctx, int steps, int *x_variant, int *y_variant, double *ms);

/* Runs every grid to its end.  ms: device time of the run; cell_updates: the
 * sum over grids of cells x completed steps.  Upload again before another run. */
int hgrp2db_run(hgrp2db *ctx, double *ms, long long *cell_updates);

/* Per grid: the step counter k (config[5] of the one-grid run), status (0, 1
 * aborted, 2 stopped), the time reached and the completed steps.  Any output
 * may be NULL. */
int hgrp2db_results(hgrp2db *ctx, int *k, int *status, double *time_c, int *steps);

/* Final states, laid out as for hgrp2db_upload; e (total energy per mass) may be NULL.
 * hgrp2db_results and hgrp2db_download refuse a context never uploaded. */
int hgrp2db_download(hgrp2db *ctx, double *rho, double *u, double *v, double *p, double *e);

/* ------------------------------------------------------------------------- */
/* Batched interface solves (host arrays in/out; for unit parity).            */
/* Interface sides are 15 doubles: gamma, lambda_u, lambda_v, rho, u, v, p,   */
/* d_rho, d_u, d_v, d_p, t_rho, t_u, t_v, t_p.                                */
/* ------------------------------------------------------------------------- */

/* GRP_2D_flux (src/flux_calc/flux_solver.c:158-237) for the split path
 * (lambda = 0, tangential derivatives taken as zero): dir 0 = x, 1 = y.
 * out: count x 9 = check code, F_rho F_u F_v F_e, RHO_int U_int V_int P_int. */
int hgrp_grp2d_flux_batch(const double *L, const double *R, int count, double tau,
                          double eps, int dir, double *out);

/* hgrp_grp2d_flux_batch without allocating device memory: the same arguments
 * and results, the staging held in a fixed device array of 65536 interfaces
 * (19.5 MiB per GPU, kept until the process ends).  No counterpart in the
 * reference; for callers that cannot allocate (code under graph capture,
 * serving loops).  count above 65536: -1002 and a message on stderr.  Calls on
 * one GPU share its array: calls that overlap on one GPU race. */
int hgrp_grp2d_flux_batch_noalloc(const double *L, const double *R, int count, double tau,
                                  double eps, int dir, double *out);

/* GRP_2D_flux with the reference's multi-fluid members
 * (src/flux_calc/flux_solver.c:158-237 with -DMULTIFLUID_BASICS, its branch at
 * :203-235; Q1D's z_a at linear_grp_solver_Edir_Q1D.c:55-61) as the two-fluid
 * split sweeps solve it: normal x (dir 0) or y (dir 1), tangential
 * derivatives zero, each side's gamma from its volume fraction z_a (fluid a:
 * gamma_a, the reference's config[6]; fluid b: gamma_b, its config[106]), the
 * energy flux with gamma at t + tau/2 from z_a there.  Sides are 21 doubles:
 * the 15 above, then z_a d_z_a t_z_a phi d_phi t_phi (t_z_a and phi unread).
 * out: count x 10 = check code, F_rho F_u F_v F_e, RHO_int U_int V_int P_int,
 * z_a at t + tau. */
int hgrp_grp2d_flux_tf_batch(const double *L, const double *R, int count, double tau, double eps, int dir,
                             double gamma_a, double gamma_b, double *out);

/* Riemann_solver_exact (src/riemann_solver/riemann_solver_exact_Ben.c:30-211,
 * declared at src/include/riemann_solver.h:24-27).
 * in: count x 10 = gammaL gammaR u_L u_R p_L p_R c_L c_R eps tol;
 * out: count x 6 = U_star P_star CRW0 CRW1 gap Newton-iterations, gap being
 * the reference's return value (the velocity gap left at the end). */
int hgrp_riemann_batch(const double *in, int count, int N, double *out);

/* GuessP (src/riemann_solver/riemann_solver_starPU.c:20-58), the reference's
 * PVRS / two-shock starting pressure, per row of in[count*11] = DL DR UL UR PL
 * PR CL CR GammaL GammaR eps; out[count]. */
int hgrp_guessp_batch(const double *in, int count, double *out);

/* The exact Riemann solve (Riemann_solver_exact) started from GuessP instead of
 * the halved tangent intersection: at most 3 Newton steps from it, then the
 * reference's own guess, so the answer is the exact solver's to its tolerance.
 * in[count*10] = rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R eps tol (as
 * hgrp_riemann_guess_batch); out[count*6] = u* p* CRW0 CRW1 Newton-iterations
 * starting-guess. */
int hgrp_riemann_pvrs_batch(const double *in, int count, int N, double *out);

/* linear_GRP_solver_Edir_Q1D (src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554,
 * src/include/riemann_solver.h:66) with all of its arguments: atc (eps: the
 * quasi-1D GRP with its acoustic case; INFINITY: acoustic only; -0.0:
 * nonlinear only), the grid velocity of L (lambda_u, lambda_v), a gamma per
 * side, both sides' tangential derivatives.
 * out: count x 16 = wave_speed[2], D[4], U[4] (rho u v p), U_star[6]
 * (rho_star_L u_star rho_star_R p_star c_star_L c_star_R). */
int hgrp_q1d_batch(const double *L, const double *R, int count, double eps, double atc, double *out);

/* linear_GRP_solver_Edir_G2D (src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637,
 * src/include/riemann_solver.h:70), the genuinely-2D direct Eulerian GRP
 * solver as the reference builds it by default (EXACT_TANGENT_DERIVATIVE
 * undefined), with all of its arguments: atc (eps: the GRP with its acoustic
 * case; INFINITY: acoustic only; -0.0: nonlinear only), lambda_u of L and
 * lambda_v of R (the reference reads them so), a gamma per side.  Sides are
 * 21 doubles: the 15 of hgrp_q1d_batch, then z_a d_z_a t_z_a phi d_phi t_phi
 * (the reference's multi-fluid members; a single-fluid caller passes
 * 0 -0 -0 0 -0 -0, the constants its single-fluid build uses).
 * out: count x 20 = wave_speed[2], D[6], U[6] (rho u v p z_a phi), U_star[6]. */
int hgrp_g2d_batch(const double *L, const double *R, int count, double eps, double atc, double *out);

/* GRP_2D_flux (src/flux_calc/flux_solver.c:158-237, src/include/flux_calc.h:28)
 * in full: unit normal (n_x, n_y), tangential derivatives and gamma of both
 * sides (lambda is set to 0 as the reference does; the energy flux uses L's
 * gamma).  out: count x 9 as hgrp_grp2d_flux_batch (fluxes zero when the
 * check code is not 0). */
int hgrp_grp2d_flux_general_batch(const double *L, const double *R, int count, double tau, double eps,
                                  double n_x, double n_y, double *out);

/* linear_GRP_solver_Edir (src/riemann_solver/linear_grp_solver_Edir.c:33-337,
 * src/include/riemann_solver.h:60), 1D: reads gamma of L and rho u p, d_rho
 * d_u d_p of both sides.  out: count x 6 = D[3], U[3] (rho u p). */
int hgrp_edir_batch(const double *L, const double *R, int count, double eps, double atc, double *out);

/* ------------------------------------------------------------------------- */
/* Stages of the 2D step over a whole grid, for callers of the reference's    */
/* per-stage functions (host arrays in/out; the time step fuses them).        */
/* Fields are 4 arrays of m*n doubles in the order rho u v p, cell (j, i) at  */
/* j*n + i (j the x index).                                                   */
/* ------------------------------------------------------------------------- */

/* The minmod slopes of bound_cond_slope_limiter_x (dir 0,
 * src/inter_process/bound_cond_slope_limiter_x.c:83-92 with minmod_limiter_2D_x,
 * slope_limiter_2D_x.c:32-84) or _y (dir 1, with minmod_limiter, slope_limiter.c:31-83):
 * ghost_lo / ghost_hi hold the ghost values rho u v p of each line (dir 0: n
 * lines along x, ghosts left/right; dir 1: m lines along y, ghosts down/up),
 * 4 doubles per line; found != 0 selects minmod3(alpha s_L, alpha s_R, s_prev)
 * (s_prev: the previous slopes, same layout as s_out), else minmod2(s_L, s_R). */
int hgrp2d_slopes(int dir, int m, int n, const double *state, const double *ghost_lo, const double *ghost_hi,
                  const double *s_prev, int found, double alpha, double h, double *s_out);

/* The interface fluxes of flux_generator_x (dir 0, src/flux_calc/flux_generator_x.c:31-165)
 * or flux_generator_y (dir 1, flux_generator_y.c:31-165): GRP_2D_flux with atc = eps
 * at every interface.  slope_n: the slopes along dir (x-slopes for dir 0,
 * y-slopes for dir 1); slope_t: the other direction's, used when transversal
 * != 0 (negated for dir 1, as the reference does; may be NULL otherwise).
 * bfv_lo / bfv_hi: per line 12 doubles = the ghost's rho u v p, its slopes
 * along dir, its transversal slopes (the b_f_var members RHO U V P, S*, T* for
 * dir 0; RHO U V P, T*, S* for dir 1).  Out, nif = (m+1)*n (dir 0, interface
 * (j, i) at j*n + i) or m*(n+1) (dir 1, at j*(n+1) + i): F[4*nif] = F_rho F_u
 * F_v F_e, I[4*nif] = rho u v p at t + tau, codes[nif] = 0, 1 (ifvar_check
 * failed), 2 (the solver's check failed; F and I are NaN there). */
int hgrp2d_fluxes(int dir, int m, int n, const double *state, const double *slope_n, const double *slope_t,
                  const double *bfv_lo, const double *bfv_hi, double tau, double h, double gamma, double eps,
                  int transversal, double *F, double *I, int *codes);

/* pow(x[k], y[k]) as evaluated on the device (precision check of the GPU's FP64 pow). */
int hgrp_pow_batch(const double *x, const double *y, int count, double *out);

/* Streams `bytes` from one device buffer to another `reps` times (a plain copy
 * kernel, copy_probe_kernel); the calibration point for HBM byte counters and
 * the measured copy bandwidth.  ms_per_copy: average device time of one copy. */
int hgrp_copy_probe(size_t bytes, int reps, double *ms_per_copy);

/* ------------------------------------------------------------------------- */
/* Batched 1D Eulerian GRP: independent problems of uneven sizes advanced    */
/* together on one GPU, each with its own CFL time step.  Each instance runs  */
/* GRP_solver_EUL_source (src/finite_volume/grp_solver_EUL_source.c:27-334,   */
/* src/include/finite_volume.h:37) on its own cells: same limiter, ghosts,    */
/* linear_GRP_solver_Edir, per-step tau from the reconstructed states, stop   */
/* and abort rules; only the final state is kept (no plot frames).            */
/* ------------------------------------------------------------------------- */
typedef struct hgrp1d hgrp1d;

/* n_inst instances with cells[i] cells of width h[i]; config[] as the
 * reference's (1 t_all, 4 eps, 5 max steps, 6 gamma, 7 CFL, 16 fixed tau,
 * 17 boundary condition, 41 minmod alpha), shared by all instances. */
int hgrp1d_create(hgrp1d **ctx, int n_inst, const int *cells, const double *h, const double *config);
int hgrp1d_destroy(hgrp1d *ctx);
/* Initial data, instances concatenated (sum of cells[] doubles each); e may
 * be NULL (then E = u^2/2 + p/(gamma-1)/rho, src/hydrocode_1D/hydrocode.c:235). */
int hgrp1d_upload(hgrp1d *ctx, const double *rho, const double *u, const double *p, const double *e);
int hgrp1d_download(hgrp1d *ctx, double *rho, double *u, double *p, double *e);
/* Runs every instance to its end.  use_graph: launch steps as captured
 * hipGraphs (1) or one by one (0).  ms: device time; cell_updates: cells times
 * steps that updated them, summed over instances (either may be NULL). */
int hgrp1d_run(hgrp1d *ctx, int use_graph, double *ms, long long *cell_updates);
/* As hgrp1d_run, but instance i also stops at the start of a step once its
 * time has reached t_pause[i] (INFINITY: never), the test the reference makes
 * before it copies a plot frame (grp_solver_EUL_source.c:116-125); the test
 * is made after every step of this call, so a call advances each unfinished
 * instance by at least one step.  paused[i] (may be NULL): 1 where instance i
 * stopped that way (a later call continues it), 0 where it has finished. */
int hgrp1d_run_until(hgrp1d *ctx, int use_graph, const double *t_pause, int *paused);
/* Per instance: the reference's step counter k (config[5] on return), status
 * (0 normal end, 1 aborted, 2 stopped by a check), and the time reached. */
int hgrp1d_results(hgrp1d *ctx, int *steps, int *status, double *time_end);
/* Autotuner: times probe_steps steps from the uploaded data for each pair of
 * block sizes (64..512) of the interface and update kernels, keeps the
 * fastest, restores the uploaded data.  The chosen sizes are returned. */
int hgrp1d_tune(hgrp1d *ctx, int probe_steps, int *bs_iface, int *bs_update);
/* Sets the block sizes (each one of 64, 128, 256, 512) of the interface and
 * update kernels, as hgrp1d_tune would; results do not depend on them. */
int hgrp1d_set_blocks(hgrp1d *ctx, int bs_iface, int bs_update);

/* ------------------------------------------------------------------------- */
/* Learned starting guess for p* (no counterpart in the reference).  A 6-32-32-1 */
/* GELU MLP predicts log(p* / p_TR), p_TR the two-rarefaction estimate; inputs   */
/* log(rho_L/rho_R)/ln1e3, log(p_L/p_R)/ln1e4, (u_R-u_L)/(5c_L), log(c_R/c_L)/5, */
/* gamma_L-1.4, gamma_R-1.4.  Master parameters: W1[32][6] b1[32] W2[32][32]   */
/* b2[32] w3[32] b3 (FP32); the device runs BF16 weights with FP32 accumulation */
/* on MFMA.  Newton takes at most 3 steps from the guess, then restarts from   */
/* the reference's guess.                                                       */
/* ------------------------------------------------------------------------- */
#define HGRP_MLP_NP 1313

/* count samples of the training distribution (left state (1,0,1), rho_L/rho_R
 * log-uniform 1e-3..1e3, p_L/p_R log-uniform 1e-4..1e4, (u_R-u_L)/c_L uniform
 * -5..5; half with gamma 1.4 on both sides, half with gamma_L, gamma_R uniform
 * in 1.2..1.8; counter-based from seed), their p* from the exact solver at
 * tolerance tol with the converge rule (HGRP_NEWTON_CONVERGE): feat[count*6], target[count] = log(p* / p_TR), ok[count];
 * states[count*9] (may be NULL) = rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R p*. */
int hgrp_mlp_data(unsigned long long seed, int count, double tol, float *feat, float *target,
                  double *states, int *ok);

/* Mean squared error over count samples and its gradient (HGRP_MLP_NP floats)
 * with respect to the master parameters, from the GPU forward/backward kernels. */
int hgrp_mlp_loss_grad(const float *params, const float *feat, const float *target, int count,
                       double *loss, float *grad);

/* Network outputs for count samples (MFMA inference path). */
int hgrp_mlp_infer(const float *params, const float *feat, int count, float *out);

/* Trains on n_samples fresh samples (hgrp_mlp_data(seed, ...)) for `epochs`
 * passes of Adam in minibatches of `batch` (a multiple of 64), learning rate lr
 * with cosine decay to lr/100.  params: initial values in, trained values out.
 * epoch_loss[epochs] (may be NULL): mean squared error of each pass. */
int hgrp_mlp_train(unsigned long long seed, int n_samples, int epochs, int batch, double lr, double tol,
                   float *params, float *epoch_loss);

/* Exact Riemann solve per interface with an optional learned guess (params
 * NULL: the reference's guess) and a stopping rule (HGRP_NEWTON_*).
 * in[count*10]: rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R eps tol;
 * out[count*6]: u* p* CRW[0] CRW[1] Newton-iterations starting-guess (-1 where
 * the reference's guess was used). */
int hgrp_riemann_guess_batch(const double *in, int count, int N, int rule, const float *params, double *out);

/* Name of the device the library runs on (empty if none). */
int hgrp_device_name(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
