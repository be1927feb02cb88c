/*
 * CPU oracle for the per-RK-stage right-hand-side pipeline of the
 * shock-turbulence solver (triply periodic, single fluid).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference Fortran algorithm, operation for operation, used as the
 * checker for the HIP product path and as the timed CPU baseline.
 * Nothing in the product library links or calls it.
 *
 * Layout: every field is Fortran-ordered F(My,Mz,Mx): y is unit stride,
 * z has stride My, x has stride My*Mz (src/main.f:28, src/header:7-10).
 * Multi-field arrays append the field index last: Q(My,Mz,Mx,5) with
 * Q = [rho*u, rho*v, rho*w, rho, rho*E].
 *
 * Parity pinning of each function is listed in the header of
 * st_oracle.c.
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;                  /* grid points (single rank: Mx=Nx ...) */
    double period_x, period_y, period_z;
    double x_start, y_start, z_start;
    double Re, Pr, gamma, vis_exp, bulk_ratio;
    int hyper_param;                 /* 1: artificial transport on */
    double cfl_cal;
    int d2_mode;                     /* 0: consistent FDL3DI D2, 1: as shipped */
    int rk_scheme;                   /* 0: CC5 then LDDRK6; 1: LDDRK5 then LDDRK6; 2: CC5 only */
    int mutate_q;                    /* 1: replay RHS_yz's in-place momentum round trip */
    int nonper_x;                    /* 1: x non-periodic (src/derivCDopt6_nonperiodic1d.f,
                                        src/filters_nonperiodic1d.f); y, z periodic */
} or_params;

typedef struct or_ctx or_ctx;

or_ctx *or_create(const or_params *p);
void or_destroy(or_ctx *c);
int or_threads(int n);               /* set OpenMP threads (n<=0: leave), return current */

/* grid metrics (src/grid.f:15-112, src/define_periodic_sf.f:42-44) */
void or_metrics(const or_ctx *c, double *out /* sx1[3], sx2[3], DX, DY, DZ */);

/* compact operators, dir 0=x 1=y 2=z (src/derivCDopt6_periodic.f) */
void or_dfd(or_ctx *c, int dir, const double *F, double *dF, int iadd, double cadd);
void or_d2fd2(or_ctx *c, int dir, const double *F, double *dF, double *d2F, int iadd);
/* filters (src/filters_periodic.f) */
void or_compact_filter(or_ctx *c, const double *F, double *F0, int nf);
void or_gaussian_filter(or_ctx *c, const double *F, double *Fg, int nf);
void or_filter_dir(or_ctx *c, int dir, const double *F, double *F0);
void or_gauss_dir(or_ctx *c, int dir, const double *F, double *Fg);
/* LHS band rows (A..E = lower-2 .. upper+2, 5*n doubles) of the non-periodic
   x operators: op 0 first derivative, 1 second derivative, 2 compact filter */
int or_np_lhs(const or_ctx *c, int op, double *rows5n);
/* pentadiagonal_serial (src/pentadiagonal.f90:319-374) on nlines lines of
   length n (line-major R) with the given band rows */
void or_np_band_solve(int n, const double *rows5n, double *R, long nlines);

/* cyclic band solve with constant bands (beta, alpha, 1, alpha, beta),
 * lines of length n with element stride 1, nlines consecutive lines
 * (src/pentadiagonal.f90:381-504) */
void or_band_solve(int n, double alpha, double beta, double *R, long nlines);

/* physics (src/rhsv_periodic_sf.f, src/hyper_dynamic.f, src/hyper_rhs.f) */
void or_rhs(or_ctx *c, double *Q, double *Qrhs);
void or_compute_hyper(or_ctx *c, const double *Q);
void or_hyper_rhs(or_ctx *c, const double *Q, double *Qrhs);
void or_get_grad(const or_ctx *c, double *out12);     /* /GRAD/ after or_rhs */
void or_get_hyper(const or_ctx *c, double *out3);     /* beta, visc, k after hyper */
void or_set_hyper(or_ctx *c, const double *in3);

/* time integration (src/lddrk4_periodic_sf.f) */
void or_fix_pressure(or_ctx *c, double *Q);
/* points Fix_Pressure has clamped on this context (test probe) */
long or_clamps(const or_ctx *c);
void or_compute_cfl(or_ctx *c, const double *Q, double *stab, double *resv);
double or_compute_tstep(or_ctx *c, const double *Q, double cfl_resv, int step);
void or_rk_stage(or_ctx *c, double *Q, double *QW, double *dQ, double alpha, double beta, double dt);
/* Sponge_Forcing (src/sponge1d.f) applied after Hyper_RHS in every stage;
   amp = 0 turns it off */
void or_set_sponge(or_ctx *c, double amp, double power, double frac, const double *ref_left, const double *ref_right);
void or_lddrk4(or_ctx *c, double *Q, double dt, double *time, int *step);
/* one march-mode-1 iteration of src/main.f:129-136: CFL, dt, LDDRK4 */
double or_march(or_ctx *c, double *Q, double *time, int *step);
/* Compute_Grad (src/grad.f:13-115): velocity part of /GRAD/, MaxDil, Enstrophy */
void or_compute_grad(or_ctx *c, const double *Q, double *maxdil, double *enstrophy);
/* MomX, MomY, MomZ, KE, Mass, Energy, Entropy (src/lddrk4_periodic_sf.f:153-302) */
void or_diagnostics(or_ctx *c, const double *Q, double *out7);

/* RK coefficient tables (src/header:253-296, init/header:208-236) */
int or_rk_coeffs(int scheme, int loop, double *alpha, double *beta);

#ifdef __cplusplus
}
#endif
#endif
