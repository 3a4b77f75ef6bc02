/*
 * var_struc.h -- the cell-variable layout and globals shared with the
 * reference's 2D program (single-fluid build).
 *
 * This is a layout contract, not new design: the member order of
 * struct cell_var_stru, struct flu_var, struct i_f_var and struct b_f_var is
 * that of src/include/var_struc.h:48-57, 61-74, 116-142, 146-151 built without
 * MULTIFLUID_BASICS (the reference's own program build never defines it), so
 * that CV arrays built by either program are interchangeable.  config[] has
 * the index meaning of doc/config.csv.
 */
#ifndef VARSTRUC_H
#define VARSTRUC_H

#ifndef EPS
#define EPS 1e-9          /* default "largest value seen as zero" (config[4]) */
#endif
#ifndef N_CONF
#define N_CONF 400        /* entries of config[] */
#endif

extern double config[];

/* initial fluid variables read from data_in (row-major file order) */
typedef struct flu_var {
    double *RHO, *U, *V, *P;
} Fluid_Variable;

/* structured-grid cell variables: X[j][i], j = x index (row pointer), i = y index */
typedef struct cell_var_stru {
    double **E;                           /* specific total energy            */
    double **RHO, **U, **V, **P;          /* density, velocity, pressure      */
    double *d_rho, *d_u, *d_v, *d_p;      /* 1D slopes                        */
    double **s_rho, **s_u, **s_v, **s_p;  /* x-slopes                         */
    double **t_rho, **t_u, **t_v, **t_p;  /* y-slopes                         */
    double **rhoIx, **uIx, **vIx, **pIx;  /* x-interface values at t_{n+1}    */
    double **rhoIy, **uIy, **vIy, **pIy;  /* y-interface values at t_{n+1}    */
    double **F_rho, **F_e, **F_u, **F_v;  /* x fluxes                         */
    double **G_rho, **G_e, **G_u, **G_v;  /* y fluxes                         */
} Cell_Variable_Structured;

/* one side of a cell interface (the argument of the per-interface solvers) */
typedef struct i_f_var {
    double n_x, n_y;                      /* unit normal of the interface     */
    double delta_x, delta_y;              /* interface centre to cell centre  */
    double length;                        /* interface length                 */
    double lambda_u, lambda_v;            /* grid velocity                    */
    double gamma;                         /* specific heat ratio              */
    double RHO, P, U, V;                  /* primitive values at t_n          */
    double RHO_int, P_int, U_int, V_int;  /* interface values at t_{n+1}      */
    double F_rho, F_e, F_u, F_v;          /* fluxes at t_{n+1/2}              */
    double U_rho, U_e, U_u, U_v;          /* conservative values at t_n       */
    double d_rho, d_e, d_u, d_v, d_p;     /* normal derivatives               */
    double t_rho, t_e, t_u, t_v, t_p;     /* tangential derivatives           */
} Interface_Fluid_Variable;

/* ghost-cell values and slopes at a boundary */
typedef struct b_f_var {
    double H;
    double RHO, P, U, V;
    double SRHO, SP, SU, SV;
    double TRHO, TP, TU, TV;
} Boundary_Fluid_Variable;

#endif
