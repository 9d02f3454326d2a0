/*
 * CPU oracle for the fractional-step dynamical-core step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the parity tests,
 * the smoke check and the cpu_baseline leg of the benchmark call it; the
 * product library (libaeolus_hip.so) never links, loads or calls it.
 *
 * It restates, operation for operation and in the same evaluation order,
 * the reference's CPU operators (Julia, badulion/Aeolus.jl), each function
 * citing the reference file:line it follows. FP64 throughout, compiled with
 * -ffp-contract=off (Julia does not contract a*b+c into an FMA).
 *
 * Storage layout (the reference's, src/Grids.jl:98-140): an array of a
 * field is (Nx+2H) x (Ny+2H) x (Nz+2H), column-major, x fastest; the
 * interior is 1-based H+1..H+N on every axis.
 */
#ifndef AEOLUS_ORACLE_H
#define AEOLUS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid metrics (src/FiniteVolume.jl:36-186). stretched = 0: UniformZ with
 * dz; stretched = 1: LevelZ vectors over the storage levels (1-based k in
 * the reference; here entry k-1). idzf has sz+1 entries. */
typedef struct {
    int Nx, Ny, Nz, H;
    double dx, dy, dz;
    int stretched;
    const double *idzc, *idzf, *wlo, *whi, *alo, *ahi;
} orc_grid;

/* Field locations (src/Grids.jl:60-83). */
enum { ORC_CC = 0, ORC_FX = 1, ORC_FY = 2, ORC_FZ = 3 };
/* Per-axis boundary kinds: Periodic, NoFlux (mirror), ImpermeableWall. */
enum { ORC_PERIODIC = 0, ORC_NOFLUX = 1, ORC_WALL = 2 };
/* Face reconstructions (src/AdvectionSchemes.jl:505-599, 1704-1775). */
enum {
    ORC_CENTERED = 0,   /* Centered2 / advect_centered! / momentum_advect! */
    ORC_UPWIND1 = 1,    /* _UpwindFirst                                      */
    ORC_WENO5 = 2,      /* FiniteVolume._weno5_left (advect_weno5!)          */
    ORC_WENO_JS5 = 3,   /* weno_js5_left                                     */
    ORC_WENO_Z5 = 4,    /* _weno_z5_left                                     */
    ORC_WENO_Z7 = 5,    /* _weno_z7_left                                     */
    ORC_TENO5 = 6,      /* _teno5_left (q = 6, cut-off orc_set_teno_cutoff)  */
    ORC_TENO8 = 7       /* _teno8_left (q = 6)                               */
};

double orc_weno_z5_left(double v1, double v2, double v3, double v4, double v5, int p);
double orc_weno_z7_left(double v1, double v2, double v3, double v4, double v5,
                        double v6, double v7, int p);
double orc_weno_js5_left(double v1, double v2, double v3, double v4, double v5, double eps);
double orc_weno5_classic_left(double fm2, double fm1, double fc, double fp1, double fp2);

void orc_fill_halos(double *a, const orc_grid *g, int loc, int bcx, int bcy, int bcz);

void orc_advect_scalar(double *tend, const double *u, const double *v, const double *w,
                       const double *phi, const orc_grid *g, int scheme, int p);
/* OpenMP threads of the following calls, and the current team size */
void orc_set_threads(int n);
int orc_get_threads(void);
/* TENO's cut-off C_T for the following calls (default 1e-5, TENO(N; C_T)) */
void orc_set_teno_cutoff(double C_T);
double orc_teno5_left(double v1, double v2, double v3, double v4, double v5, double C_T);
double orc_teno8_left(double v1, double v2, double v3, double v4, double v5, double v6, double v7, double v8,
                      double C_T);
/* advect_variance_conserving! (src/AdvectionSchemes.jl:1428-1446), regular grids */
void orc_advect_variance_conserving(double *tend, const double *u, const double *v, const double *w,
                                    const double *phi, const orc_grid *g);
/* advect_positive! (src/AdvectionSchemes.jl:1135-1160): limited fluxes so a
 * forward-Euler update with the same dt keeps ϕ ≥ 0; use_bcs = 0 wraps the
 * limiter's halo periodically by one layer (the reference's bcs = nothing) */
void orc_advect_positive(double *tend, const double *u, const double *v, const double *w, const double *phi,
                         const orc_grid *g, int scheme, int p, double dt, int use_bcs, int bcx, int bcy, int bcz,
                         int accumulate, double *ratio);
void orc_advect_momentum(double *tu, double *tv, double *tw, const double *u, const double *v,
                         const double *w, const orc_grid *g, int scheme, int p);
void orc_laplacian(double *out, const double *f, const orc_grid *g, int loc);
void orc_divergence_into(double *buf, const double *u, const double *v, const double *w,
                         const orc_grid *g);
double orc_max_divergence(const double *u, const double *v, const double *w, const orc_grid *g);
double orc_max_adv_rate(const double *u, const double *v, const double *w, const orc_grid *g);

/* Canonical plane reduction (src/Backends.jl:703-750) and ordered_sum. */
double orc_interior_sum(const double *a, int sx, int sy, int i1, int n_i, int j1, int n_j,
                        int k1, int n_k);

/* Direct solvers. */
void *orc_fft_create(int Nx, int Ny, int Nz, double dx, double dy, double dz);
void orc_fft_solve(void *h, double *phi_interior, const double *src);
void *orc_column_create(int Nx, int Ny, int Nz, double dx, double dy, int bcx, int bcy,
                        const double *lower, const double *upper, const double *hweight,
                        const double *measure, const double *inv_coupling, const double *gauge);
void orc_column_solve(void *h, double *phi_interior, const double *src);
void orc_solver_destroy(void *h);

/* The composed SSP-RK3 projection step. */
typedef struct {
    int mom_scheme, scal_scheme, weno_p;
    double nu, theta0, g_accel, flux_bot, flux_top;
    int bcx, bcy, bcz;     /* velocity BCs */
    int tbcx, tbcy, tbcz;  /* theta BCs    */
    int solver;            /* 0: FFT (triply periodic); 1: column solver */
} orc_step_params;

void *orc_stepper_create(const orc_grid *g, const orc_step_params *sp, void *solver);
void orc_stepper_destroy(void *h);
void orc_step_rk3(void *h, double *u, double *v, double *w, double *theta, double *p, double dt);
void orc_step_euler(void *h, double *u, double *v, double *w, double *theta, double *p, double dt);

/* Multigrid and MG-preconditioned CG (src/Multigrid.jl, src/Backends.jl:1461-1591). */
void *orc_mg_create(int nx, int ny, int nz, double dx, double dy, double dz, int bcx, int bcy,
                    int bcz, const double *cx, const double *cy, const double *cz,
                    const double *rhs_scale, int nu1, int nu2, int min_size, int smoother,
                    double omega);
void orc_mg_destroy(void *h);
int orc_mg_levels(void *h);
/* Plain multigrid cycles (mg_solve!). Returns cycles; *rel = last relative residual. */
int orc_mg_solve(void *h, double *x, const double *rhs, double tol, int maxcycles, double *rel);
/* MGPreconditionedCG solve!: returns status (0 converged, 1 maxit, 2 breakdown, 3 zero_rhs). */
int orc_mgpcg_solve(void *h, double *x, const double *rhs, double tol, int maxit, int *iters,
                    double *resid, double *bnorm);
/* Apply the fine operator to an interior-sized array. */
void orc_mg_apply(void *h, double *out, const double *x);
/* Dense-free PCG on the 7-point operator with identity preconditioner. */

/* Cut-cell geometry from a column heightmap (src/CutCell.jl:213-520). */
void orc_cutcell_heightmap(const orc_grid *g, const double *height_storage, int nsub,
                           double vol_min, double *vol, double *vol_eff, double *ax, double *ay,
                           double *az);

#ifdef __cplusplus
}
#endif
#endif
