/*
 * hydrocode_dropin.h -- entry points with the reference's names and
 * signatures, implemented by this project (lib/libhydrocode2d.so, sources in
 * hydrocode-fusion_amd/host/).
 *
 * The solver is a drop-in for the reference's dimension-split 2D Eulerian GRP
 * path: a program that calls it as src/hydrocode_2D/hydrocode.c:264-265 does
 * gets the same CV frames, config[5], N_plot and time_plot back, computed on
 * the GPU through the C ABI of include/hydrocode_grp.h.  The I/O routines keep
 * the reference's data_in / data_out formats.
 */
#ifndef HYDROCODE_DROPIN_H
#define HYDROCODE_DROPIN_H

#include <stdio.h>
#include "var_struc.h"

#ifdef __cplusplus
typedef bool hgrp_bool;     /* C's _Bool and C++'s bool share size and ABI */
extern "C" {
#else
typedef _Bool hgrp_bool;
#endif

/* replaces GRP_solver_2D_split_EUL_source, src/include/finite_volume.h:48-49
 * (src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379) */
void GRP_solver_2D_split_EUL_source(const int m, const int n, struct cell_var_stru *CV, double **X, double **Y,
                                    double *cpu_time, const char *problem, int N_T, int *N_plot,
                                    double time_plot[]);

/* replaces GRP_solver_2D_EUL_source, src/include/finite_volume.h:43-44
 * (src/finite_volume/grp_solver_2D_EUL_source.c:75-330): the unsplit scheme */
void GRP_solver_2D_EUL_source(const int m, const int n, struct cell_var_stru *CV, double **X, double **Y,
                              double *cpu_time, const char *problem, int N_T, int *N_plot, double time_plot[]);

/* replaces GRP_solver_EUL_source, src/include/finite_volume.h:37
 * (src/finite_volume/grp_solver_EUL_source.c:27-334): the 1D scheme; CV.X[k]
 * are the plot frames of m cells (host/lib_grp_solver_1D.c) */
void GRP_solver_EUL_source(const int m, struct cell_var_stru CV, double *cpu_time, int *N_plot, double time_plot[]);

/* The 2D step's stages, over the whole grid on the GPU (host/lib_stages.c):
 * the same CV members, b_f_var arrays and return values as the reference's;
 * at an interface whose solver check fails the fluxes stored are NaN. */
/* replaces bound_cond_slope_limiter_x, src/include/inter_process.h:51-52
 * (src/inter_process/bound_cond_slope_limiter_x.c:27-129) */
hgrp_bool bound_cond_slope_limiter_x(const int m, const int n, const int nt, struct cell_var_stru *CV,
                                 struct b_f_var *bfv_L, struct b_f_var *bfv_R, struct b_f_var *bfv_D,
                                 struct b_f_var *bfv_U, hgrp_bool find_bound_x, const hgrp_bool Slope, const double t_c);
/* replaces bound_cond_slope_limiter_y, src/include/inter_process.h:56-57
 * (src/inter_process/bound_cond_slope_limiter_y.c:27-129) */
hgrp_bool bound_cond_slope_limiter_y(const int m, const int n, const int nt, struct cell_var_stru *CV,
                                 struct b_f_var *bfv_L, struct b_f_var *bfv_R, struct b_f_var *bfv_D,
                                 struct b_f_var *bfv_U, hgrp_bool find_bound_y, const hgrp_bool Slope, const double t_c);
/* replaces flux_generator_x, src/include/flux_calc.h:16-17 (src/flux_calc/flux_generator_x.c:31-165) */
int flux_generator_x(const int m, const int n, const int nt, const double tau, struct cell_var_stru *CV,
                     struct b_f_var *bfv_L, struct b_f_var *bfv_R, const hgrp_bool Transversal);
/* replaces flux_generator_y, src/include/flux_calc.h:21-22 (src/flux_calc/flux_generator_y.c:31-165) */
int flux_generator_y(const int m, const int n, const int nt, const double tau, struct cell_var_stru *CV,
                     struct b_f_var *bfv_D, struct b_f_var *bfv_U, const hgrp_bool Transversal);

/* The per-interface solvers, one GPU thread per call (host/lib_units.c).
 * Single-fluid members only (the reference's program build has no
 * MULTIFLUID_BASICS): D and U of the Q1D solver carry rho u v p.  A failed
 * device call is reported on stderr and yields NaN outputs (GRP_2D_flux: -1). */
/* replaces Riemann_solver_exact, src/include/riemann_solver.h:24-27
 * (src/riemann_solver/riemann_solver_exact_Ben.c:30-211) */
double Riemann_solver_exact(double *U_star, double *P_star, const double gammaL, const double gammaR,
                            const double u_L, const double u_R, const double p_L, const double p_R,
                            const double c_L, const double c_R, hgrp_bool *CRW, const double eps, const double tol,
                            int N);
/* replaces linear_GRP_solver_Edir, src/include/riemann_solver.h:60
 * (src/riemann_solver/linear_grp_solver_Edir.c:33-337) */
void linear_GRP_solver_Edir(double *D, double *U, const struct i_f_var *ifv_L, const struct i_f_var *ifv_R,
                            const double eps, const double atc);
/* replaces linear_GRP_solver_Edir_Q1D, src/include/riemann_solver.h:66
 * (src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554) */
void linear_GRP_solver_Edir_Q1D(double *wave_speed, double *D, double *U, double *U_star,
                                const struct i_f_var *ifv_L, const struct i_f_var *ifv_R, const double eps,
                                const double atc);
/* replaces linear_GRP_solver_Edir_G2D, src/include/riemann_solver.h:70
 * (src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637, default build
 * without EXACT_TANGENT_DERIVATIVE); D and U receive six entries each
 * (rho u v p z_a phi, the last two the single-fluid build's constants) */
void linear_GRP_solver_Edir_G2D(double *wave_speed, double *D, double *U, double *U_star,
                                const struct i_f_var *ifv_L, const struct i_f_var *ifv_R, const double eps,
                                const double atc);
/* replaces GRP_2D_flux, src/include/flux_calc.h:28 (src/flux_calc/flux_solver.c:158-237);
 * eps is config[4] */
int GRP_2D_flux(struct i_f_var *ifv, struct i_f_var *ifv_R, const double tau);

/* replaces the 2D program's file I/O, src/include/file_io.h:13-54 */
void example_io(const char *example, char *add_mkdir, const int i_or_o);        /* io_control.c:40-115   */
int flu_var_count(FILE *fp, const char *add);                                    /* io_control.c:124-153  */
int flu_var_count_line(FILE *fp, const char *add, int *n_x);                     /* io_control.c:163-210  */
int flu_var_read(FILE *fp, double *U, const int num);                            /* io_control.c:221-261  */
int time_plot_read(const char *add_in, const int N_max, int *N_plot, double **time_plot); /* io_control.c:288-335 */
void arg_preprocess(const int argc_least, const int argc, char *argv[], char *scheme);     /* terminal_io.c:26-98 */
void configurate(const char *name);                                              /* config_handle.c:208-243 */
void config_write(const char *add_out, const double *cpu_time, const char *name);/* config_handle.c:253-300 */
struct flu_var initialize_2D(const char *name, int *N, int *N_plot, double **time_plot); /* file_2D_in.c:119-179 */
void file_2D_write(const int n_x, const int n_y, const int N, const struct cell_var_stru CV[],
                   double **X, double **Y, const double *cpu_time, const char *problem,
                   const double time_plot[]);                                    /* file_2D_out.c:55-88   */
void file_2D_write_POINT_TEC(const int n_x, const int n_y, const int N, const struct cell_var_stru CV[],
                             double **X, double **Y, const double *cpu_time, const char *problem,
                             const double time_plot[]);                          /* file_2D_out.c:103-151 */

/* Full-precision frames, written beside the reference's outputs (no counterpart
 * in the reference, whose .dat/.tec files carry %.10g and whose 2D HDF5 output
 * is float32).  FLU_VAR.f64: magic "HGRPF64\0", int32 version 1, n_x, n_y, N,
 * double time_plot[N], then per frame RHO U V P E as n_x*n_y doubles each
 * (j outer, i inner).  Returns 0 on success. */
int file_2D_write_f64(const char *file, const int n_x, const int n_y, const int N,
                      const struct cell_var_stru CV[], const double time_plot[]);
/* Reads such a file; *time_plot (N) and *frames (N*5*n_x*n_y) are malloc'd.
 * Returns 0, 1 (cannot open) or 2 (not a frame file / short). */
int file_2D_read_f64(const char *file, int *n_x, int *n_y, int *N, double **time_plot, double **frames);

/* replaces src/tools/sys_pro.c:37-100 */
void DispPro(const double pro, const int step);
int CreateDir(const char *pPath);

#ifdef __cplusplus
}
#endif
#endif
