/*
 * st_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the periodic
 * single-fluid per-RK-stage pipeline of the shock-turbulence solver.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use this file, and only as the checker or the timed CPU baseline.
 *
 * It restates the reference Fortran operation for operation (same
 * expression association, same elimination order in the band solver),
 * compiled with -ffp-contract=off so no fused multiply-adds appear.
 * Each routine cites the reference file:line it follows (paths relative
 * to the reference repository root).
 *
 * How each part is pinned (the tests under tests/ carry the checks):
 *   band solve (pentadiagonal_periodic_serial)  dense numpy solve, n=5..16
 *   D1, D2, compact filter, Gaussian            closed-form symbols on
 *                                               sinusoids (to roundoff),
 *                                               6th-order refinement,
 *                                               D2(const)=0, T(pi)=0, G(0)=1
 *   RHS, /GRAD/, Compute_coeff1,                a second restatement from
 *   Compute_hyper, Hyper_RHS, one stage          the reference text with every
 *                                               periodic compact operator
 *                                               applied through its exact
 *                                               discrete symbol in Fourier
 *                                               space (tests/spectral_ref.py):
 *                                               agreement to 1e-15..8e-15
 *                                               (3e-13 for a stage, through
 *                                               the filter's band solve) on
 *                                               multi-mode states with the
 *                                               artificial transport on; plus
 *                                               closed forms of the sensor,
 *                                               uniform state -> 0, mass,
 *                                               momentum and energy
 *                                               conservation, axis-swap
 *                                               symmetry, a manufactured
 *                                               Euler solution at 6th order,
 *                                               and a planted-mistake test
 *   non-periodic closure rows, serial solve,    polynomial exactness of each
 *   mirrored Gaussian                           closure row (and failure one
 *                                               degree above), dense LU
 *                                               solves, numpy's reflect-padded
 *                                               convolution
 *   libm at the boundary (tanh, sqrt, pow)      parity unpinned (no pinned
 *                                               runtime in the reference)
 *   RK stage / step                             4th-order convergence on a
 *                                               linear ODE through the same
 *                                               coefficient tables
 *   Fix_Pressure, Compute_CFL, Compute_Tstep    hand algebra, forced clamp,
 *                                               closed-form dt on a uniform
 *                                               state
 */
#include "st_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- */
/* coefficients: src/header:191-203, src/LES_header:107-113,153-157 */
/* ---------------------------------------------------------------- */
static const double a1_pade6 = 14.0 / 18.0;
static const double b1_pade6 = 1.0 / 36.0;
static const double c1_pade6 = 0.0;
static const double alpha1_pade6 = 1.0 / 3.0;
static const double a2_pade6 = 12.0 / 11.0;
static const double b2_pade6 = 3.0 / 44.0;
static const double alpha2_pade6 = 2.0 / 11.0;

static const double alpha_filt4 = 6.6624e-01;
static const double beta_filt4 = 1.6688e-01;
static const double a0_filt4 = 9.9965e-01;
static const double a1_filt4 = 0.66652e+00;
static const double a2_filt4 = 0.16674e+00;
static const double a3_filt4 = 4.0e-05;
static const double a4_filt4 = -5.0e-06;

static const double a0_filt_gauss = 3565.0 / 10368.0;
static const double a1_filt_gauss = 3091.0 / 12960.0;
static const double a2_filt_gauss = 1997.0 / 25920.0;
static const double a3_filt_gauss = 149.0 / 12960.0;
static const double a4_filt_gauss = 107.0 / 103680.0;

/* CFL constants, src/header:237-240 */
static const double k_max = 2.478;
static const double k_hug = 2.199;

/* ---------------------------------------------------------------- */
/* RK tables: CC5 src/header:284-296; LDDRK6 src/header:267-281     */
/* (identical to init/header:222-236); LDDRK5 src/header:253-265     */
/* ---------------------------------------------------------------- */
static const double cc5_alpha[5] = {0.0, -4.801594388477952e-01, -1.404247195200064e+00,
                                    -2.016477077503357e+00, -1.056444269767207e+00};
static const double cc5_beta[5] = {1.028639988104959e-01, 7.408540575766301e-01,
                                   7.426530946683799e-01, 4.694937902357905e-01,
                                   1.881733382887932e-01};
static const double ld6_alpha[6] = {0.0,
                                    -4.4127377153877382565e-1,
                                    -1.073982008079781868e0,
                                    -1.7063570791256758809e0,
                                    -2.7979293162682443056e0,
                                    -4.0913537120919160454e0};
static const double ld6_beta[6] = {1.1584888181285561688e-1, 3.7287699051652864918e-1,
                                   7.3795368921435295698e-1, 5.7981109366311039538e-1,
                                   1.031284991300145194e0, 1.5e-1};
static const double ld5_alpha[5] = {0.0, -6.0512264332862261228e-1, -2.04375640234761394333e0,
                                    -0.74069990637544192841e0, -4.42317651302968168941e0};
static const double ld5_beta[5] = {2.6874543888713438496e-1, 8.014706973220802933e-1,
                                   5.0515704269422722538e-1, 5.6235680379000296407e-1,
                                   5.900655127758823335e-2};

int or_rk_coeffs(int scheme, int loop, double *alpha, double *beta)
{
    const double *a = NULL, *b = NULL;
    int n = 0;
    if (loop == 0) {
        if (scheme == 1) { a = ld5_alpha; b = ld5_beta; n = 5; }
        else { a = cc5_alpha; b = cc5_beta; n = 5; }
    } else if (loop == 1 && scheme != 2) {
        a = ld6_alpha; b = ld6_beta; n = 6;
    }
    for (int i = 0; i < n; i++) {
        if (alpha) alpha[i] = a[i];
        if (beta) beta[i] = b[i];
    }
    return n;
}

/* ---------------------------------------------------------------- */
/* cyclic band solver replay: src/pentadiagonal.f90:381-504          */
/* The band values are the same for every line, so the elimination   */
/* multipliers and border columns are computed once here with the    */
/* reference's own update sequence; band_apply then replays exactly  */
/* the operations the reference performs on the right-hand side.     */
/* ---------------------------------------------------------------- */
typedef struct {
    int n;
    double *f1, *f2;   /* forward:  R(i+1) -= R(i)*f1[i], R(i+2) -= R(i)*f2[i] */
    double *g1, *g2;   /* backward: R(i-1) -= R(i)*g1[i], R(i-2) -= R(i)*g2[i] */
    double o1, o2, o3, kk;
    double *s1, *s2, *C;
} band_t;

static band_t *band_new(int n, double alpha, double beta)
{
    band_t *b = (band_t *)calloc(1, sizeof(band_t));
    b->n = n;
    /* 1-based arrays of length n+1 */
    double *A = calloc(n + 2, sizeof(double)), *B = calloc(n + 2, sizeof(double));
    double *C = calloc(n + 2, sizeof(double)), *D = calloc(n + 2, sizeof(double));
    double *E = calloc(n + 2, sizeof(double));
    double *s1 = calloc(n + 2, sizeof(double)), *s2 = calloc(n + 2, sizeof(double));
    b->f1 = calloc(n + 2, sizeof(double)); b->f2 = calloc(n + 2, sizeof(double));
    b->g1 = calloc(n + 2, sizeof(double)); b->g2 = calloc(n + 2, sizeof(double));
    for (int i = 1; i <= n; i++) { A[i] = beta; B[i] = alpha; C[i] = 1.0; D[i] = alpha; E[i] = beta; }
    if (n == 4) {               /* :417-420 */
        for (int i = 1; i <= n; i++) { A[i] = A[i] + E[i]; E[i] = 0.0; }
    }
    /* :423-434 */
    s1[1] = A[1];
    s1[n - 3] = s1[n - 3] + E[n - 3];
    s1[n - 2] = D[n - 2];
    s1[n - 1] = C[n - 1];
    s1[n] = B[n];
    s2[1] = B[1];
    s2[2] = A[2];
    s2[n - 2] = s2[n - 2] + E[n - 2];
    s2[n - 1] = D[n - 1];
    s2[n] = C[n];
    /* forward elimination :437-453 */
    for (int i = 1; i <= n - 2; i++) {
        double c = B[i + 1] / C[i];
        b->f1[i] = c;
        C[i + 1] = C[i + 1] - D[i] * c;
        D[i + 1] = D[i + 1] - E[i] * c;
        s1[i + 1] = s1[i + 1] - s1[i] * c;
        s2[i + 1] = s2[i + 1] - s2[i] * c;
        c = A[i + 2] / C[i];
        b->f2[i] = c;
        B[i + 2] = B[i + 2] - D[i] * c;
        C[i + 2] = C[i + 2] - E[i] * c;
        s1[i + 2] = s1[i + 2] - s1[i] * c;
        s2[i + 2] = s2[i + 2] - s2[i] * c;
    }
    /* backward elimination :456-474 */
    for (int i = n - 2; i >= 3; i--) {
        double c = D[i - 1] / C[i];
        b->g1[i] = c;
        s1[i - 1] = s1[i - 1] - s1[i] * c;
        s2[i - 1] = s2[i - 1] - s2[i] * c;
        c = E[i - 2] / C[i];
        b->g2[i] = c;
        s1[i - 2] = s1[i - 2] - s1[i] * c;
        s2[i - 2] = s2[i - 2] - s2[i] * c;
    }
    {
        double c = D[1] / C[2];
        b->g1[2] = c;
        s1[1] = s1[1] - s1[2] * c;
        s2[1] = s2[1] - s2[2] * c;
    }
    /* oddball :477-490 */
    b->o1 = E[n - 1] / C[1];
    s1[n - 1] = s1[n - 1] - s1[1] * b->o1;
    s2[n - 1] = s2[n - 1] - s2[1] * b->o1;
    b->o2 = D[n] / C[1];
    s1[n] = s1[n] - s1[1] * b->o2;
    s2[n] = s2[n] - s2[1] * b->o2;
    b->o3 = E[n] / C[2];
    s1[n] = s1[n] - s1[2] * b->o3;
    s2[n] = s2[n] - s2[2] * b->o3;
    /* corner :493-495 */
    b->kk = s1[n] / s1[n - 1];
    s2[n] = s2[n] - s2[n - 1] * b->kk;
    b->s1 = s1; b->s2 = s2; b->C = C;
    free(A); free(B); free(D); free(E);
    return b;
}

static void band_free(band_t *b)
{
    if (!b) return;
    free(b->f1); free(b->f2); free(b->g1); free(b->g2);
    free(b->s1); free(b->s2); free(b->C); free(b);
}

/* R is a 1-based line R[1..n] (pass pointer to element 0 slot). */
static void band_apply(const band_t *b, double *R)
{
    const int n = b->n;
    for (int i = 1; i <= n - 2; i++) {
        R[i + 1] = R[i + 1] - R[i] * b->f1[i];
        if (b->f2[i] != 0.0) R[i + 2] = R[i + 2] - R[i] * b->f2[i];
    }
    for (int i = n - 2; i >= 3; i--) {
        R[i - 1] = R[i - 1] - R[i] * b->g1[i];
        if (b->g2[i] != 0.0) R[i - 2] = R[i - 2] - R[i] * b->g2[i];
    }
    R[1] = R[1] - R[2] * b->g1[2];
    if (b->o1 != 0.0) R[n - 1] = R[n - 1] - R[1] * b->o1;
    R[n] = R[n] - R[1] * b->o2;
    if (b->o3 != 0.0) R[n] = R[n] - R[2] * b->o3;
    R[n] = R[n] - R[n - 1] * b->kk;
    R[n] = R[n] / b->s2[n];
    R[n - 1] = (R[n - 1] - b->s2[n - 1] * R[n]) / b->s1[n - 1];
    for (int i = n - 2; i >= 1; i--)
        R[i] = (R[i] - b->s1[i] * R[n - 1] - b->s2[i] * R[n]) / b->C[i];
}

void or_band_solve(int n, double alpha, double beta, double *R, long nlines)
{
    band_t *b = band_new(n, alpha, beta);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < nlines; l++) band_apply(b, R + l * n - 1);
    band_free(b);
}

/* ---------------------------------------------------------------- */
/* context                                                          */
/* ---------------------------------------------------------------- */
struct or_ctx {
    or_params p;
    int n[3];                   /* Mx, My, Mz */
    long npts;
    double sx1[3], sx2[3], D[3]; /* metric terms and DX, DY, DZ */
    double re_inv, pr_inv;
    double c2_d2, d2_d2;        /* D2 +-3 weight and centre weight */
    band_t *b1[3], *b2[3], *bf[3];
    double *np_rows[3];         /* non-periodic x: LHS rows of D1, D2, filter (5*Mx each) */
    double sp_amp, sp_pow, sp_frac, sp_ref[10];   /* sponge (amp 0: off) */
    double *grad;               /* /GRAD/: dux_dx duy_dx duz_dx dux_dy duy_dy duz_dy
                                   dux_dz duy_dz duz_dz dTdx dTdy dTdz */
    double *hyp;                /* /HYPER/: beta_hyper, visc_hyper, k_hyper */
    double *arena;              /* persistent scratch fields (first-touched in parallel) */
    long arena_fields, arena_top;
    long clamps;                /* points Fix_Pressure clamped since creation (a probe for tests) */
};

/* Scratch fields come from a persistent arena used as a stack, so repeated
   calls do not page-fault fresh allocations. */
static double *ws_get(or_ctx *c, long nfields)
{
    if (c->arena_top + nfields > c->arena_fields) abort();
    double *p = c->arena + (size_t)c->arena_top * c->npts;
    c->arena_top += nfields;
    return p;
}

static void pset(double *dst, double v, size_t n)
{
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < n; l++) dst[l] = v;
}

static void pcopy(double *dst, const double *src, size_t n)
{
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < n; l++) dst[l] = src[l];
}

static int dir_n(const or_ctx *c, int d) { return c->n[d]; }

/* Lines along direction d: count, base index, element stride. */
static long dir_lot(const or_ctx *c, int d)
{
    return c->npts / c->n[d];
}

static void dir_line(const or_ctx *c, int d, long line, long *base, long *stride)
{
    const long My = c->n[1], Mz = c->n[2];
    if (d == 0) {                 /* x: lines (j,k) */
        *base = line;             /* j + My*k */
        *stride = My * Mz;
    } else if (d == 1) {          /* y: lines (k,i) contiguous */
        *base = line * My;
        *stride = 1;
    } else {                      /* z: lines (j,i) */
        long j = line % My, i = line / My;
        *base = j + My * Mz * i;
        *stride = My;
    }
}

int or_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

or_ctx *or_create(const or_params *p)
{
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->p = *p;
    c->n[0] = p->nx; c->n[1] = p->ny; c->n[2] = p->nz;
    c->npts = (long)p->nx * p->ny * p->nz;
    const double period[3] = {p->period_x, p->period_y, p->period_z};
    const double start[3] = {p->x_start, p->y_start, p->z_start};
    for (int d = 0; d < 3; d++) {
        const int N = c->n[d];
        /* src/define_periodic_sf.f:42-44; a non-periodic x spans its
           period with both end points (src/define_nonperiodic1d_sf.f:42) */
        const double leng = (d == 0 && p->nonper_x) ? period[d] : period[d] * (double)(N - 1) / (double)N;
        /* src/grid.f:31-32 */
        c->sx1[d] = 1.0 / leng;
        c->sx2[d] = c->sx1[d] * c->sx1[d];
        /* src/grid.f:132-135 then DX = gxcor(2) - gxcor(1) */
        const double g1 = start[d] + leng * (double)(0) / (double)(N - 1);
        const double g2 = start[d] + leng * (double)(1) / (double)(N - 1);
        c->D[d] = g2 - g1;
        c->b1[d] = band_new(N, alpha1_pade6, 0.0);
        c->b2[d] = band_new(N, alpha2_pade6, 0.0);
        c->bf[d] = band_new(N, alpha_filt4, beta_filt4);
    }
    /* src/define_periodic_sf.f:179-189 */
    c->re_inv = p->Re > 0.0 ? 1.0 / p->Re : 0.0;
    c->pr_inv = p->Pr > 0.0 ? 1.0 / p->Pr : 0.0;
    /* D2 mapping: src/header:198-203 names c2 = 2(a2+b2), d2 = 0; the
       periodic routines use c2 on the +-3 points and d2 on the centre
       (src/derivCDopt6_periodic.f:326-329).  Mode 0 maps the FDL3DI
       centre weight consistently: +-3 weight 0, centre 2(a2+b2). */
    if (p->d2_mode == 1) {
        c->c2_d2 = 2.0 * (a2_pade6 + b2_pade6);
        c->d2_d2 = 0.0;
    } else {
        c->c2_d2 = 0.0;
        c->d2_d2 = 2.0 * (a2_pade6 + b2_pade6);
    }
    if (p->nonper_x)
        for (int op = 0; op < 3; op++) {
            c->np_rows[op] = malloc(sizeof(double) * 5 * c->n[0]);
            or_np_lhs(c, op, c->np_rows[op]);
        }
    c->grad = malloc(sizeof(double) * 12 * c->npts);
    c->hyp = malloc(sizeof(double) * 3 * c->npts);
    pset(c->grad, 0.0, (size_t)12 * c->npts);
    pset(c->hyp, 0.0, (size_t)3 * c->npts);
    c->arena_fields = 96;
    c->arena_top = 0;
    c->arena = malloc(sizeof(double) * c->arena_fields * c->npts);
    pset(c->arena, 0.0, (size_t)c->arena_fields * c->npts);
    return c;
}

void or_destroy(or_ctx *c)
{
    for (int op = 0; op < 3; op++) free(c->np_rows[op]);
    if (!c) return;
    for (int d = 0; d < 3; d++) { band_free(c->b1[d]); band_free(c->b2[d]); band_free(c->bf[d]); }
    free(c->grad); free(c->hyp); free(c->arena); free(c);
}

void or_metrics(const or_ctx *c, double *out)
{
    for (int d = 0; d < 3; d++) { out[d] = c->sx1[d]; out[3 + d] = c->sx2[d]; out[6 + d] = c->D[d]; }
}


/* ---------------------------------------------------------------- */
/* DFDX/DFDY/DFDZ: src/derivCDopt6_periodic.f:164-277,400-512,631-744 */
/* ---------------------------------------------------------------- */
static void np_dfdx(or_ctx *c, const double *F, double *dF, int iadd, double cadd);
static void np_d2fdx2(or_ctx *c, const double *F, double *dF, double *d2F, int iadd);
static void np_filter_x(or_ctx *c, const double *F, double *F0);
static void np_gauss_x(or_ctx *c, const double *F, double *Fg);

void or_dfd(or_ctx *c, int d, const double *F, double *dF, int iadd, double cadd)
{
    if (d == 0 && c->p.nonper_x) { np_dfdx(c, F, dF, iadd, cadd); return; }
    const int n = dir_n(c, d);
    const long lot = dir_lot(c, d);
    /* fac = cadd * sx(1) * (Nx - 1)   (:181, :417, :648) */
    const double fac = cadd * c->sx1[d] * (double)(n - 1);
    const band_t *b = c->b1[d];
#pragma omp parallel
    {
        double *fb = malloc(sizeof(double) * (n + 8)), *W = malloc(sizeof(double) * (n + 1));
        double *f = fb + 4;
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, d, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 1; i <= 4; i++) { f[-i] = f[n - i]; f[n - 1 + i] = f[i - 1]; }
            for (int i = 0; i < n; i++) {
                W[i + 1] = fac * (c1_pade6 * (f[i + 3] - f[i - 3])
                                + b1_pade6 * (f[i + 2] - f[i - 2])
                                + a1_pade6 * (f[i + 1] - f[i - 1]));
            }
            band_apply(b, W);
            if (iadd == 0)
                for (int i = 0; i < n; i++) dF[base + i * st] = W[i + 1];
            else
                for (int i = 0; i < n; i++) dF[base + i * st] = dF[base + i * st] + W[i + 1];
        }
        free(fb); free(W);
    }
}

/* D2FDX2/D2FDY2/D2FDZ2: :281-397, :515-628, :747-862.  With iadd=1 the
   first derivative is produced by a DFDX(F,dF,0,1) call (:381). */
void or_d2fd2(or_ctx *c, int d, const double *F, double *dF, double *d2F, int iadd)
{
    if (d == 0 && c->p.nonper_x) { np_d2fdx2(c, F, dF, d2F, iadd); return; }
    const int n = dir_n(c, d);
    const long lot = dir_lot(c, d);
    /* fac = ((Nx - 1) * (Nx - 1)) * sx(2)   (:298) */
    const double fac = (double)((n - 1) * (n - 1)) * c->sx2[d];
    const double c2 = c->c2_d2, dd = c->d2_d2;
    const band_t *b = c->b2[d];
    if (iadd == 1) or_dfd(c, d, F, dF, 0, 1.0);
#pragma omp parallel
    {
        double *fb = malloc(sizeof(double) * (n + 8)), *W = malloc(sizeof(double) * (n + 1));
        double *f = fb + 4;
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, d, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 1; i <= 4; i++) { f[-i] = f[n - i]; f[n - 1 + i] = f[i - 1]; }
            for (int i = 0; i < n; i++) {
                W[i + 1] = fac * (c2 * (f[i + 3] + f[i - 3])
                                + b2_pade6 * (f[i + 2] + f[i - 2])
                                + a2_pade6 * (f[i + 1] + f[i - 1])
                                - dd * f[i]);
            }
            band_apply(b, W);
            for (int i = 0; i < n; i++) d2F[base + i * st] = W[i + 1];
        }
        free(fb); free(W);
    }
}

/* ---------------------------------------------------------------- */
/* filters: src/filters_periodic.f                                   */
/* ---------------------------------------------------------------- */
/* Compact_Filter_X/Y/Z :115-251, :254-390, :393-528 (F0 may alias F) */
void or_filter_dir(or_ctx *c, int d, const double *F, double *F0)
{
    if (d == 0 && c->p.nonper_x) { np_filter_x(c, F, F0); return; }
    const int n = dir_n(c, d);
    const long lot = dir_lot(c, d);
    const band_t *b = c->bf[d];
#pragma omp parallel
    {
        double *fb = malloc(sizeof(double) * (n + 8)), *W = malloc(sizeof(double) * (n + 1));
        double *f = fb + 4;
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, d, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 1; i <= 4; i++) { f[-i] = f[n - i]; f[n - 1 + i] = f[i - 1]; }
            for (int i = 0; i < n; i++) {
                W[i + 1] = a4_filt4 * (f[i + 4] + f[i - 4])
                         + a3_filt4 * (f[i + 3] + f[i - 3])
                         + a2_filt4 * (f[i + 2] + f[i - 2])
                         + a1_filt4 * (f[i + 1] + f[i - 1])
                         + a0_filt4 * f[i];
            }
            band_apply(b, W);
            for (int i = 0; i < n; i++) F0[base + i * st] = W[i + 1];
        }
        free(fb); free(W);
    }
}

/* Compact_Filter :90-111: X into F0, then Y and Z in place */
void or_compact_filter(or_ctx *c, const double *F, double *F0, int nf)
{
    for (int L = 0; L < nf; L++) {
        const double *f = F + (size_t)L * c->npts;
        double *f0 = F0 + (size_t)L * c->npts;
        or_filter_dir(c, 0, f, f0);
        or_filter_dir(c, 1, f0, f0);
        or_filter_dir(c, 2, f0, f0);
    }
}

/* Gaussian_Filter_X/Y/Z :565-678, :680-791, :793-906 */
void or_gauss_dir(or_ctx *c, int d, const double *F, double *Fg)
{
    if (d == 0 && c->p.nonper_x) { np_gauss_x(c, F, Fg); return; }
    const int n = dir_n(c, d);
    const long lot = dir_lot(c, d);
#pragma omp parallel
    {
        double *fb = malloc(sizeof(double) * (n + 8));
        double *f = fb + 4;
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, d, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 1; i <= 4; i++) { f[-i] = f[n - i]; f[n - 1 + i] = f[i - 1]; }
            for (int i = 0; i < n; i++) {
                Fg[base + i * st] = a4_filt_gauss * (f[i + 4] + f[i - 4])
                                  + a3_filt_gauss * (f[i + 3] + f[i - 3])
                                  + a2_filt_gauss * (f[i + 2] + f[i - 2])
                                  + a1_filt_gauss * (f[i + 1] + f[i - 1])
                                  + a0_filt_gauss * f[i];
            }
        }
        free(fb);
    }
}

/* ---------------------------------------------------------------- */
/* non-periodic x operators                                          */
/* ---------------------------------------------------------------- */
/* closures: src/header:170-175 */
static const double closure1 = 5.0 / 9.0, closure2 = 1.0 / 18.0;
static const double closure3 = 1.0 / 36.0, closure4 = 7.0 / 9.0;
static const double closure5 = 3.0 / 44.0, closure6 = 12.0 / 11.0;
/* Visbal & Gaitonde near-boundary filter: src/LES_header:118-148 */
static const double bndy_alpha_filt = 0.45;
#define BA bndy_alpha_filt
static const double bc1[5] = {15.0 / 16.0, 1.0 / 4.0, -3.0 / 8.0, 1.0 / 4.0, -1.0 / 16.0};
static const double bc2[5] = {1.0 / 16.0 + (7.0 / 8.0) * BA, 3.0 / 4.0 + (1.0 / 2.0) * BA,
                              3.0 / 8.0 + (1.0 / 4.0) * BA, -1.0 / 4.0 + (1.0 / 2.0) * BA,
                              1.0 / 16.0 - (1.0 / 8.0) * BA};
static const double bc3_0 = 5.0 / 8.0 + (3.0 / 4.0) * BA;
static const double bc3_1 = (1.0 / 2.0 + BA) / 2.0;
static const double bc3_2 = (-1.0 / 8.0 + BA / 4.0) / 2.0;
#undef BA

/* Rows 1..n of the LHS bands (A=lower-2, B=lower-1, C=diag, D=upper+1,
   E=upper+2), stored as rows[5*i + (0..4)].
   op 0: Setup_Derivatives first derivative, src/derivCDopt6_nonperiodic1d.f:41-158
   op 1: second derivative :160-357
   op 2: Init_Filters src/filters_nonperiodic1d.f:45-215 */
int or_np_lhs(const or_ctx *c, int op, double *r)
{
    const int n = c->n[0];
    double a, b;
    if (op == 0) { a = 0.0; b = alpha1_pade6; }          /* beta1_pade6 = 0 (src/header:193-197) */
    else if (op == 1) { a = 0.0; b = alpha2_pade6; }     /* beta2_pade6 = 0 (src/header:198-203) */
    else { a = beta_filt4; b = alpha_filt4; }
    for (int i = 0; i < n; i++) {
        r[5 * i + 0] = a; r[5 * i + 1] = b; r[5 * i + 2] = 1.0; r[5 * i + 3] = b; r[5 * i + 4] = a;
    }
#define ROW(i, A, B, C, D, E) do { double *q = r + 5 * (i); q[0] = A; q[1] = B; q[2] = C; q[3] = D; q[4] = E; } while (0)
    if (op == 0) {
        ROW(0, 0.0, 0.0, 1.0, 2.0, 0.0);
        ROW(1, 0.0, 1.0 / 6.0, 1.0, 1.0 / 2.0, 0.0);
        ROW(2, 0.0, 1.0 / 3.0, 1.0, 1.0 / 3.0, 0.0);
        ROW(n - 1, 0.0, 2.0, 1.0, 0.0, 0.0);
        ROW(n - 2, 0.0, 1.0 / 2.0, 1.0, 1.0 / 6.0, 0.0);
        ROW(n - 3, 0.0, 1.0 / 3.0, 1.0, 1.0 / 3.0, 0.0);
    } else if (op == 1) {
        ROW(0, 0.0, 0.0, 1.0, 11.0, 0.0);
        ROW(1, 0.0, 1.0e-1, 1.0, -7.0 / 20.0, 0.0);
        ROW(2, 0.0, 2.0 / 11.0, 1.0, 2.0 / 11.0, 0.0);
        ROW(n - 1, 0.0, 11.0, 1.0, 0.0, 0.0);
        ROW(n - 2, 0.0, -7.0 / 20.0, 1.0, 1.0e-1, 0.0);
        ROW(n - 3, 0.0, 2.0 / 11.0, 1.0, 2.0 / 11.0, 0.0);
    } else {
        ROW(0, 0.0, 0.0, 1.0, 0.0, 0.0);
        ROW(1, 0.0, bndy_alpha_filt, 1.0, bndy_alpha_filt, 0.0);
        ROW(2, 0.0, bndy_alpha_filt, 1.0, bndy_alpha_filt, 0.0);
        ROW(3, beta_filt4, alpha_filt4, 1.0, alpha_filt4, beta_filt4);
        ROW(n - 1, 0.0, 0.0, 1.0, 0.0, 0.0);
        ROW(n - 2, 0.0, bndy_alpha_filt, 1.0, bndy_alpha_filt, 0.0);
        ROW(n - 3, 0.0, bndy_alpha_filt, 1.0, bndy_alpha_filt, 0.0);
        ROW(n - 4, beta_filt4, alpha_filt4, 1.0, alpha_filt4, beta_filt4);
    }
#undef ROW
    return 0;
}

/* pentadiagonal_serial, src/pentadiagonal.f90:319-374, for one line:
   the bands are copied per call (the reference copies them, :42) and
   tiny(1.0) (single precision) guards each pivot. */
static void np_solve_line(int n, const double *rows, double *R, double *w)
{
    double *A = w, *B = w + n, *C = w + 2 * n, *D = w + 3 * n, *E = w + 4 * n;
    const double tiny1 = 1.17549435082228750797e-38;    /* tiny(1.0) */
    for (int i = 0; i < n; i++) {
        A[i] = rows[5 * i]; B[i] = rows[5 * i + 1]; C[i] = rows[5 * i + 2];
        D[i] = rows[5 * i + 3]; E[i] = rows[5 * i + 4];
    }
    /* 0-based: reference row i+1 is index i */
    for (int i = 0; i < n - 2; i++) {
        double cst = B[i + 1] / (C[i] + tiny1);
        C[i + 1] = C[i + 1] - D[i] * cst;
        D[i + 1] = D[i + 1] - E[i] * cst;
        R[i + 1] = R[i + 1] - R[i] * cst;
        cst = A[i + 2] / (C[i] + tiny1);
        B[i + 2] = B[i + 2] - D[i] * cst;
        C[i + 2] = C[i + 2] - E[i] * cst;
        R[i + 2] = R[i + 2] - R[i] * cst;
    }
    {
        const double cst = B[n - 1] / (C[n - 2] + tiny1);
        C[n - 1] = C[n - 1] - D[n - 2] * cst;
        R[n - 1] = R[n - 1] - R[n - 2] * cst;
    }
    R[n - 1] = R[n - 1] / (C[n - 1] + tiny1);
    R[n - 2] = (R[n - 2] - D[n - 2] * R[n - 1]) / (C[n - 2] + tiny1);
    for (int i = n - 3; i >= 0; i--)
        R[i] = (R[i] - D[i] * R[i + 1] - E[i] * R[i + 2]) / (C[i] + tiny1);
}

void or_np_band_solve(int n, const double *rows, double *R, long nlines)
{
#pragma omp parallel
    {
        double *w = malloc(sizeof(double) * 5 * n);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; l++) np_solve_line(n, rows, R + (size_t)l * n, w);
        free(w);
    }
}

/* DFDX src/derivCDopt6_nonperiodic1d.f:393-541 (single rank: both end
   closures, :449-506) */
static void np_dfdx(or_ctx *c, const double *F, double *dF, int iadd, double cadd)
{
    const int n = c->n[0];
    const long lot = dir_lot(c, 0);
    const double fac = cadd * c->sx1[0] * (double)(n - 1);
#pragma omp parallel
    {
        double *f = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * 5 * n);
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, 0, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 3; i < n - 3; i++)
                W[i] = fac * (c1_pade6 * (f[i + 3] - f[i - 3]) + b1_pade6 * (f[i + 2] - f[i - 2])
                              + a1_pade6 * (f[i + 1] - f[i - 1]));
            W[2] = fac * (closure3 * (f[4] - f[0]) + closure4 * (f[3] - f[1]));
            W[1] = fac * (-closure1 * f[0] - 5.0e-1 * f[1] + f[2] + closure2 * f[3]);
            W[0] = fac * (-2.5 * f[0] + 2.0 * f[1] + 5.0e-1 * f[2]);
            W[n - 3] = fac * (closure3 * (f[n - 1] - f[n - 5]) + closure4 * (f[n - 2] - f[n - 4]));
            W[n - 2] = fac * (closure1 * f[n - 1] + 5.0e-1 * f[n - 2] - f[n - 3] - closure2 * f[n - 4]);
            W[n - 1] = fac * (2.5 * f[n - 1] - 2.0 * f[n - 2] - 5.0e-1 * f[n - 3]);
            np_solve_line(n, c->np_rows[0], W, w);
            if (iadd == 0)
                for (int i = 0; i < n; i++) dF[base + i * st] = W[i];
            else
                for (int i = 0; i < n; i++) dF[base + i * st] = dF[base + i * st] + W[i];
        }
        free(f); free(W); free(w);
    }
}

/* D2FDX2 :544-705; interior with the D2 mapping of the periodic routines
   (c2 on +-3, d2 at the centre) */
static void np_d2fdx2(or_ctx *c, const double *F, double *dF, double *d2F, int iadd)
{
    const int n = c->n[0];
    const long lot = dir_lot(c, 0);
    const double fac = (double)((n - 1) * (n - 1)) * c->sx2[0];
    const double c2 = c->c2_d2, dd = c->d2_d2;
#pragma omp parallel
    {
        double *f = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * 5 * n);
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, 0, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 3; i < n - 3; i++)
                W[i] = fac * (c2 * (f[i + 3] + f[i - 3]) + b2_pade6 * (f[i + 2] + f[i - 2])
                              + a2_pade6 * (f[i + 1] + f[i - 1]) - dd * f[i]);
            W[2] = fac * (closure5 * (f[0] - 2.0 * f[2] + f[4]) + closure6 * (f[1] - 2.0 * f[2] + f[3]));
            W[1] = fac * (1.2375 * f[0] - 3.0 * f[1] + 2.325 * f[2] - 6.0e-1 * f[3] + 3.75e-2 * f[4]);
            W[0] = fac * (1.3e1 * f[0] - 2.7e1 * f[1] + 1.5e1 * f[2] - f[3]);
            W[n - 3] = fac * (closure5 * (f[n - 1] - 2.0 * f[n - 3] + f[n - 5])
                              + closure6 * (f[n - 2] - 2.0 * f[n - 3] + f[n - 4]));
            W[n - 2] = fac * (1.2375 * f[n - 1] - 3.0 * f[n - 2] + 2.325 * f[n - 3] - 6.0e-1 * f[n - 4]
                              + 3.75e-2 * f[n - 5]);
            W[n - 1] = fac * (1.3e1 * f[n - 1] - 2.7e1 * f[n - 2] + 1.5e1 * f[n - 3] - f[n - 4]);
            np_solve_line(n, c->np_rows[1], W, w);
            for (int i = 0; i < n; i++) d2F[base + i * st] = W[i];
        }
        free(f); free(W); free(w);
    }
    if (iadd == 1) np_dfdx(c, F, dF, 0, 1.0);          /* :697 */
}

/* Compact_Filter_X src/filters_nonperiodic1d.f:257-456 */
static void np_filter_x(or_ctx *c, const double *F, double *F0)
{
    const int n = c->n[0];
    const long lot = dir_lot(c, 0);
#pragma omp parallel
    {
        double *f = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * 5 * n);
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, 0, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int i = 4; i < n - 4; i++)
                W[i] = a4_filt4 * (f[i + 4] + f[i - 4]) + a3_filt4 * (f[i + 3] + f[i - 3])
                     + a2_filt4 * (f[i + 2] + f[i - 2]) + a1_filt4 * (f[i + 1] + f[i - 1]) + a0_filt4 * f[i];
            W[0] = bc1[0] * f[0] + bc1[1] * f[1] + bc1[2] * f[2] + bc1[3] * f[3] + bc1[4] * f[4];
            W[1] = bc2[0] * f[0] + bc2[1] * f[1] + bc2[2] * f[2] + bc2[3] * f[3] + bc2[4] * f[4];
            W[2] = bc3_2 * f[4] + bc3_1 * f[3] + bc3_0 * f[2] + bc3_1 * f[1] + bc3_2 * f[0];
            {
                const double a0 = 11.0 / 16.0 + (5.0 / 8.0) * alpha_filt4 - (5.0 / 8.0) * beta_filt4;
                const double a1 = (15.0 / 32.0 + (17.0 / 16.0) * alpha_filt4 + (15.0 / 16.0) * beta_filt4) / 2.0;
                const double a2 = (-3.0 / 16.0 + (3.0 / 8.0) * alpha_filt4 + (13.0 / 8.0) * beta_filt4) / 2.0;
                const double a3 = (1.0 / 32.0 - alpha_filt4 / 16.0 + beta_filt4 / 16.0) / 2.0;
                W[3] = a3 * f[6] + a2 * f[5] + a1 * f[4] + a0 * f[3] + a1 * f[2] + a2 * f[1] + a3 * f[0];
                W[n - 4] = a3 * f[n - 7] + a2 * f[n - 6] + a1 * f[n - 5] + a0 * f[n - 4] + a1 * f[n - 3]
                         + a2 * f[n - 2] + a3 * f[n - 1];
            }
            W[n - 1] = bc1[0] * f[n - 1] + bc1[1] * f[n - 2] + bc1[2] * f[n - 3] + bc1[3] * f[n - 4] + bc1[4] * f[n - 5];
            W[n - 2] = bc2[0] * f[n - 1] + bc2[1] * f[n - 2] + bc2[2] * f[n - 3] + bc2[3] * f[n - 4] + bc2[4] * f[n - 5];
            W[n - 3] = bc3_2 * f[n - 5] + bc3_1 * f[n - 4] + bc3_0 * f[n - 3] + bc3_1 * f[n - 2] + bc3_2 * f[n - 1];
            np_solve_line(n, c->np_rows[2], W, w);
            for (int i = 0; i < n; i++) F0[base + i * st] = W[i];
        }
        free(f); free(W); free(w);
    }
}

/* Gaussian_Filter_X src/filters_nonperiodic1d.f:769-943: symmetry about
   the end points, f(1-m) = f(1+m) and f(N+m) = f(N-m) */
static void np_gauss_x(or_ctx *c, const double *F, double *Fg)
{
    const int n = c->n[0];
    const long lot = dir_lot(c, 0);
#pragma omp parallel
    {
        double *fb = malloc(sizeof(double) * (n + 8));
        double *f = fb + 4;
#pragma omp for schedule(static)
        for (long l = 0; l < lot; l++) {
            long base, st;
            dir_line(c, 0, l, &base, &st);
            for (int i = 0; i < n; i++) f[i] = F[base + i * st];
            for (int m = 1; m <= 4; m++) { f[-m] = f[m]; f[n - 1 + m] = f[n - 1 - m]; }
            for (int i = 0; i < n; i++) {
                double g;
                if (i == 0 || i == n - 1) {
                    const int s = i == 0 ? 1 : -1;
                    g = a4_filt_gauss * (2.0 * f[i + 4 * s]) + a3_filt_gauss * (2.0 * f[i + 3 * s])
                      + a2_filt_gauss * (2.0 * f[i + 2 * s]) + a1_filt_gauss * (2.0 * f[i + s])
                      + a0_filt_gauss * f[i];
                } else {
                    g = a4_filt_gauss * (f[i + 4] + f[i - 4]) + a3_filt_gauss * (f[i + 3] + f[i - 3])
                      + a2_filt_gauss * (f[i + 2] + f[i - 2]) + a1_filt_gauss * (f[i + 1] + f[i - 1])
                      + a0_filt_gauss * f[i];
                }
                Fg[base + i * st] = g;
            }
        }
        free(fb);
    }
}

/* Gaussian_Filter :540-560: X into Fg, Y into a temporary, Z into Fg */
void or_gaussian_filter(or_ctx *c, const double *F, double *Fg, int nf)
{
    const long mark = c->arena_top;
    double *tmp = ws_get(c, 1);
    for (int L = 0; L < nf; L++) {
        const double *f = F + (size_t)L * c->npts;
        double *fg = Fg + (size_t)L * c->npts;
        or_gauss_dir(c, 0, f, fg);
        or_gauss_dir(c, 1, fg, tmp);
        or_gauss_dir(c, 2, tmp, fg);
    }
    c->arena_top = mark;
}

/* ---------------------------------------------------------------- */
/* RHS: src/rhsv_periodic_sf.f                                       */
/* ---------------------------------------------------------------- */
#define FLD(a, k) ((a) + (size_t)(k) * N)
#define PFOR _Pragma("omp parallel for schedule(static)") for (long l = 0; l < N; l++)

/* RHS_x :36-461.  Q here is the caller's copy Ux (modified in place). */
static void rhs_x(or_ctx *c, double *Q, double *Qcx, double *Qv, double *W6)
{
    const long N = c->npts;
    const double gamma = c->p.gamma, vis_exp = c->p.vis_exp, bulk = c->p.bulk_ratio;
    const double Re_inv = c->re_inv, Pr_inv = c->pr_inv;
    double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    double *P = FLD(W6, 0), *Vs = FLD(W6, 1), *T = FLD(W6, 2);
    double *FF = FLD(W6, 3), *Flux = FLD(W6, 4), *dF = FLD(W6, 5);
    double *G = c->grad;
    double fac1 = gamma - 1.0;
    double fac2 = gamma / fac1;

    /* :92-104 */
    PFOR {
        Vs[l] = 1.0 / Q4[l];
        P[l] = fac1 * (Q5[l] - 5.0e-1 * Vs[l] * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
        T[l] = fac2 * Vs[l] * P[l];
        dF[l] = Q1[l] * Vs[l];
    }
    PFOR FF[l] = dF[l];                                      /* :110-116 */
    PFOR Flux[l] = Q4[l] * FF[l];                            /* :120-126 */
    or_dfd(c, 0, Flux, FLD(Qcx, 3), 1, -1.0);                /* :132 */
    PFOR Flux[l] = (Q5[l] + P[l]) * FF[l];                   /* :135-141 */
    or_dfd(c, 0, Flux, FLD(Qcx, 4), 1, -1.0);                /* :147 */
    PFOR FF[l] = dF[l];                                      /* :152-158 */
    PFOR Flux[l] = Q1[l] * FF[l];                            /* :163-169 */
    or_dfd(c, 0, Flux, FLD(Qcx, 0), 1, -5.0e-1);             /* :175 */
    or_dfd(c, 0, P, FLD(Qcx, 0), 1, -1.0);                   /* :181 */
    PFOR Flux[l] = Q2[l] * FF[l];                            /* :186-192 */
    or_dfd(c, 0, Flux, FLD(Qcx, 1), 1, -5.0e-1);             /* :198 */
    PFOR Flux[l] = Q3[l] * FF[l];                            /* :203-209 */
    or_dfd(c, 0, Flux, FLD(Qcx, 2), 1, -5.0e-1);             /* :215 */
    or_dfd(c, 0, Q1, FF, 0, 1.0);                            /* :221 */
    fac2 = vis_exp - 1.0;                                    /* :237 */
    PFOR {                                                   /* :239-250 */
        Q1[l] = dF[l];
        Q2[l] = Q2[l] * Vs[l];
        Q3[l] = Q3[l] * Vs[l];
        Q5[l] = pow(fac1 * T[l], vis_exp) * Re_inv;
        Flux[l] = pow(fac1 * T[l], fac2) * Re_inv * vis_exp * fac1;
    }
    PFOR {                                                   /* :266-276 */
        const double f2 = -5.0e-1 * FF[l];
        FLD(Qcx, 0)[l] = FLD(Qcx, 0)[l] + f2 * Q1[l];
        FLD(Qcx, 1)[l] = FLD(Qcx, 1)[l] + f2 * Q2[l];
        FLD(Qcx, 2)[l] = FLD(Qcx, 2)[l] + f2 * Q3[l];
    }
    or_d2fd2(c, 0, T, FF, dF, 1);                            /* :288 */
    pcopy(FLD(G, 9), FF, N);               /* :290 dTdx */
    PFOR {                                                   /* :296-306 */
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Pr_inv * (Q5[l] * dF[l] + Flux[l] * FF[l] * FF[l]);
    }
    or_d2fd2(c, 0, Q1, Flux, dF, 1);                         /* :316 */
    pcopy(FLD(G, 0), Flux, N);             /* :318 dux_dx */
    fac1 = 4.0 / 3.0 + bulk;                                 /* :328 */
    PFOR {                                                   /* :330-340 */
        const double f2 = fac1 * Q5[l] * dF[l];
        FLD(Qv, 0)[l] = FLD(Qv, 0)[l] + f2;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q1[l] * f2;
    }
    or_d2fd2(c, 0, Q2, T, dF, 1);                            /* :350 */
    pcopy(FLD(G, 1), T, N);                /* :352 duy_dx */
    PFOR {                                                   /* :368-379 */
        const double f2 = Q5[l] * dF[l];
        FLD(Qv, 1)[l] = FLD(Qv, 1)[l] + f2;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q2[l] * f2;
        Q2[l] = T[l];
    }
    or_d2fd2(c, 0, Q3, Vs, dF, 1);                           /* :389 */
    pcopy(FLD(G, 2), Vs, N);               /* :391 duz_dx */
    PFOR {                                                   /* :407-418 */
        const double f2 = Q5[l] * dF[l];
        FLD(Qv, 2)[l] = FLD(Qv, 2)[l] + f2;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q3[l] * f2;
        Q3[l] = Vs[l];
    }
    PFOR {                                                   /* :444-457 */
        const double f2 = -5.0e-1 * Q4[l] * Q1[l];
        Q1[l] = Flux[l];
        Q4[l] = FF[l];
        FLD(Qcx, 0)[l] = FLD(Qcx, 0)[l] + f2 * Q1[l];
        FLD(Qcx, 1)[l] = FLD(Qcx, 1)[l] + f2 * Q2[l];
        FLD(Qcx, 2)[l] = FLD(Qcx, 2)[l] + f2 * Q3[l];
    }
}

/* RHS_yz :472-1385.  Q is the caller's state (momenta converted to
   velocities and back in place at :732-740 and :1371-1379); Ux is the
   array RHS_x left behind: (dux_dx, duy_dx, duz_dx, dTdx, mu/Re). */
static void rhs_yz(or_ctx *c, double *Q, double *Ux, double *Qcy, double *Qcz, double *Qv, double *W)
{
    const long N = c->npts;
    const double gamma = c->p.gamma, vis_exp = c->p.vis_exp, bulk = c->p.bulk_ratio;
    const double Re_inv = c->re_inv, Pr_inv = c->pr_inv;
    double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    double *U1 = FLD(Ux, 0), *U2 = FLD(Ux, 1), *U3 = FLD(Ux, 2), *U4 = FLD(Ux, 3), *U5 = FLD(Ux, 4);
    double *P = FLD(W, 0), *Vs = FLD(W, 1), *T = FLD(W, 2), *vis = FLD(W, 3);
    double *dF = FLD(W, 4), *Flux = FLD(W, 5), *Tdz = FLD(W, 6), *Udz = FLD(W, 7);
    double *dFz = FLD(W, 8), *Udx = FLD(W, 9), *Udy = FLD(W, 10), *Tdy = FLD(W, 11);
    double *G = c->grad;
    double fac1 = gamma - 1.0;
    double fac2 = gamma / fac1;
    double fac3;

    PFOR {                                                   /* :530-544 */
        Vs[l] = 1.0 / Q4[l];
        P[l] = fac1 * (Q5[l] - 5.0e-1 * Vs[l] * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
        T[l] = fac2 * Vs[l] * P[l];
        U5[l] = Q2[l] * Vs[l];
        Flux[l] = Q1[l] * U5[l];
    }
    or_dfd(c, 1, Q2, dF, 0, 1.0);                            /* :552 */
    PFOR FLD(Qcy, 3)[l] = FLD(Qcy, 3)[l] - dF[l];            /* :558-564 */
    or_dfd(c, 1, Flux, FLD(Qcy, 0), 1, -5.0e-1);             /* :570 */
    PFOR Flux[l] = Q2[l] * U5[l];                            /* :574-580 */
    or_dfd(c, 1, Flux, FLD(Qcy, 1), 1, -5.0e-1);             /* :586 */
    or_dfd(c, 1, P, FLD(Qcy, 1), 1, -1.0);                   /* :592 */
    PFOR Flux[l] = Q3[l] * U5[l];                            /* :596-602 */
    or_dfd(c, 1, Flux, FLD(Qcy, 2), 1, -5.0e-1);             /* :608 */
    PFOR Flux[l] = (Q5[l] + P[l]) * U5[l];                   /* :612-618 */
    or_dfd(c, 1, Flux, FLD(Qcy, 4), 1, -1.0);                /* :624 */
    or_dfd(c, 2, Q3, dFz, 0, 1.0);                           /* :630 */
    PFOR FLD(Qcz, 3)[l] = FLD(Qcz, 3)[l] - dFz[l];           /* :638-644 */
    PFOR U5[l] = Q3[l] * Vs[l];                              /* :648-654 */
    PFOR Flux[l] = Q1[l] * U5[l];                            /* :658-664 */
    or_dfd(c, 2, Flux, FLD(Qcz, 0), 1, -5.0e-1);             /* :670 */
    PFOR Flux[l] = Q2[l] * U5[l];                            /* :674-680 */
    or_dfd(c, 2, Flux, FLD(Qcz, 1), 1, -5.0e-1);             /* :686 */
    PFOR Flux[l] = Q3[l] * U5[l];                            /* :690-696 */
    or_dfd(c, 2, Flux, FLD(Qcz, 2), 1, -5.0e-1);             /* :702 */
    or_dfd(c, 2, P, FLD(Qcz, 2), 1, -1.0);                   /* :708 */
    PFOR Flux[l] = (Q5[l] + P[l]) * U5[l];                   /* :712-718 */
    or_dfd(c, 2, Flux, FLD(Qcz, 4), 1, -1.0);                /* :724 */
    PFOR {                                                   /* :732-740 */
        Q1[l] = Q1[l] * Vs[l];
        Q2[l] = Q2[l] * Vs[l];
        Q3[l] = Q3[l] * Vs[l];
    }
    PFOR {                                                   /* :768-778 */
        const double f2 = -5.0e-1 * dF[l];
        FLD(Qcy, 0)[l] = FLD(Qcy, 0)[l] + f2 * Q1[l];
        FLD(Qcy, 1)[l] = FLD(Qcy, 1)[l] + f2 * Q2[l];
        FLD(Qcy, 2)[l] = FLD(Qcy, 2)[l] + f2 * Q3[l];
    }
    PFOR {                                                   /* :780-790 */
        const double f2 = -5.0e-1 * dFz[l];
        FLD(Qcz, 0)[l] = FLD(Qcz, 0)[l] + f2 * Q1[l];
        FLD(Qcz, 1)[l] = FLD(Qcz, 1)[l] + f2 * Q2[l];
        FLD(Qcz, 2)[l] = FLD(Qcz, 2)[l] + f2 * Q3[l];
    }
    or_d2fd2(c, 1, T, Tdy, U5, 1);                           /* :803 */
    pcopy(FLD(G, 10), Tdy, N);             /* :805 dTdy */
    or_d2fd2(c, 2, T, Tdz, Flux, 1);                         /* :815 */
    pcopy(FLD(G, 11), Tdz, N);             /* :817 dTdz */
    fac2 = vis_exp - 1.0;                                    /* :829 */
    PFOR {                                                   /* :831-839 */
        vis[l] = pow(fac1 * T[l], vis_exp) * Re_inv;
        Udz[l] = pow(fac1 * T[l], fac2) * Re_inv * vis_exp * fac1;
    }
    PFOR {                                                   /* :853-864 */
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Pr_inv * (vis[l] * (U5[l] + Flux[l])
                        + Udz[l] * (Tdy[l] * Tdy[l] + Tdz[l] * Tdz[l]));
    }
    PFOR {                                                   /* :878-886 */
        U4[l] = Udz[l] * U4[l];
        Tdy[l] = Udz[l] * Tdy[l];
        Tdz[l] = Udz[l] * Tdz[l];
    }
    /* :892 DFDY(Q4, Udz) is overwritten at :926 before any read; omitted. */
    or_d2fd2(c, 1, Q1, P, U5, 1);                            /* :902 */
    pcopy(FLD(G, 3), P, N);                /* :904 dux_dy */
    or_d2fd2(c, 1, Q2, Flux, dF, 1);                         /* :914 */
    pcopy(FLD(G, 4), Flux, N);             /* :916 duy_dy */
    or_d2fd2(c, 1, Q3, Udz, T, 1);                           /* :926 */
    pcopy(FLD(G, 5), Udz, N);              /* :928 duz_dy */
    or_dfd(c, 2, Q1, Udx, 0, 1.0);                           /* :935 */
    pcopy(FLD(G, 6), Udx, N);              /* :937 dux_dz */
    or_dfd(c, 2, Q2, Udy, 0, 1.0);                           /* :944 */
    pcopy(FLD(G, 7), Udy, N);              /* :946 duy_dz */
    or_dfd(c, 2, Q3, dFz, 0, 1.0);                           /* :953 */
    pcopy(FLD(G, 8), dFz, N);              /* :955 duz_dz */
    PFOR {                                                   /* :983-993 */
        const double f2 = -5.0e-1 * Q4[l] * Q2[l];
        FLD(Qcy, 0)[l] = FLD(Qcy, 0)[l] + f2 * P[l];
        FLD(Qcy, 1)[l] = FLD(Qcy, 1)[l] + f2 * Flux[l];
        FLD(Qcy, 2)[l] = FLD(Qcy, 2)[l] + f2 * Udz[l];
    }
    PFOR {                                                   /* :995-1005 */
        const double f2 = -5.0e-1 * Q4[l] * Q3[l];
        FLD(Qcz, 0)[l] = FLD(Qcz, 0)[l] + f2 * Udx[l];
        FLD(Qcz, 1)[l] = FLD(Qcz, 1)[l] + f2 * Udy[l];
        FLD(Qcz, 2)[l] = FLD(Qcz, 2)[l] + f2 * dFz[l];
    }
    fac1 = bulk + 4.0 / 3.0;                                 /* :1034 */
    PFOR {                                                   /* :1036-1052 */
        const double fX = vis[l] * U5[l];
        const double fY = fac1 * vis[l] * dF[l];
        const double fZ = vis[l] * T[l];
        FLD(Qv, 0)[l] = FLD(Qv, 0)[l] + fX;
        FLD(Qv, 1)[l] = FLD(Qv, 1)[l] + fY;
        FLD(Qv, 2)[l] = FLD(Qv, 2)[l] + fZ;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q1[l] * fX + Q2[l] * fY + Q3[l] * fZ;
    }
    or_d2fd2(c, 2, Q3, dF, T, 0);                            /* :1062 */
    pcopy(dF, FLD(G, 8), N);               /* :1064 dF = duz_dz */
    or_dfd(c, 1, dF, U5, 0, 1.0);                            /* :1070 */
    fac1 = bulk - 2.0 / 3.0;                                 /* :1124-1126 */
    fac2 = bulk + 4.0 / 3.0;
    fac3 = bulk + 1.0 / 3.0;
    PFOR {                                                   /* :1128-1154 */
        const double tmp1 = U1[l];
        const double dil = tmp1 + Flux[l] + dF[l];
        const double fX = U4[l] * (2.0 * tmp1 + fac1 * dil);
        const double fY = Tdy[l] * (2.0 * Flux[l] + fac1 * dil) + fac3 * vis[l] * U5[l];
        const double fZ = Tdz[l] * (2.0 * dF[l] + fac1 * dil) + fac2 * vis[l] * T[l];
        FLD(Qv, 0)[l] = FLD(Qv, 0)[l] + fX;
        FLD(Qv, 1)[l] = FLD(Qv, 1)[l] + fY;
        FLD(Qv, 2)[l] = FLD(Qv, 2)[l] + fZ;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q1[l] * fX + Q2[l] * fY + Q3[l] * fZ
                        + vis[l] * (2.0 * (tmp1 * tmp1 + Flux[l] * Flux[l] + dF[l] * dF[l])
                                    + fac1 * dil * dil);
    }
    or_d2fd2(c, 2, Q1, T, dF, 0);                            /* :1164 */
    pcopy(T, FLD(G, 6), N);                /* :1166 T = dux_dz */
    or_d2fd2(c, 2, Q2, Flux, U5, 0);                         /* :1176 */
    pcopy(Flux, FLD(G, 7), N);             /* :1178 Flux = duy_dz */
    PFOR {                                                   /* :1252-1275 */
        const double tmp1 = U2[l] + P[l];
        const double tmp2 = U3[l] + T[l];
        const double tmp3 = Udz[l] + Flux[l];
        const double fX = Tdy[l] * tmp1 + Tdz[l] * tmp2 + vis[l] * dF[l];
        const double fY = U4[l] * tmp1 + Tdz[l] * tmp3 + vis[l] * U5[l];
        const double fZ = U4[l] * tmp2 + Tdy[l] * tmp3;
        FLD(Qv, 0)[l] = FLD(Qv, 0)[l] + fX;
        FLD(Qv, 1)[l] = FLD(Qv, 1)[l] + fY;
        FLD(Qv, 2)[l] = FLD(Qv, 2)[l] + fZ;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q1[l] * fX + Q2[l] * fY + Q3[l] * fZ
                        + vis[l] * (tmp1 * tmp1 + tmp2 * tmp2 + tmp3 * tmp3);
    }
    or_dfd(c, 1, U1, U4, 0, 1.0);                            /* :1281 */
    or_dfd(c, 2, U1, U5, 0, 1.0);                            /* :1287 */
    or_dfd(c, 1, U2, Tdy, 0, 1.0);                           /* :1293 */
    or_dfd(c, 2, U3, Tdz, 0, 1.0);                           /* :1299 */
    or_dfd(c, 1, Flux, T, 0, 1.0);                           /* :1305 */
    PFOR {                                                   /* :1347-1363 */
        const double fX = fac3 * vis[l] * (Tdy[l] + Tdz[l]);
        const double fY = fac3 * vis[l] * U4[l];
        const double fZ = fac3 * vis[l] * (U5[l] + T[l]);
        FLD(Qv, 0)[l] = FLD(Qv, 0)[l] + fX;
        FLD(Qv, 1)[l] = FLD(Qv, 1)[l] + fY;
        FLD(Qv, 2)[l] = FLD(Qv, 2)[l] + fZ;
        FLD(Qv, 4)[l] = FLD(Qv, 4)[l] + Q1[l] * fX + Q2[l] * fY + Q3[l] * fZ;
    }
    PFOR {                                                   /* :1371-1379 */
        Q1[l] = Q1[l] * Q4[l];
        Q2[l] = Q2[l] * Q4[l];
        Q3[l] = Q3[l] * Q4[l];
    }
}

/* RHS :9-25.  With mutate_q=0 the caller's Q is left untouched (the
   right-hand side is identical either way; only Q's ulps differ). */
void or_rhs(or_ctx *c, double *Q, double *Qrhs)
{
    const long N = c->npts;
    const size_t F5 = (size_t)5 * N;
    const long mark = c->arena_top;
    double *Ux = ws_get(c, 5);
    double *Qc = ws_get(c, 20);           /* Qconvx, Qconvy, Qconvz, Qvisc */
    double *W = ws_get(c, 12);
    pset(Qc, 0.0, (size_t)4 * F5);
    double *Qcx = Qc, *Qcy = Qc + F5, *Qcz = Qc + 2 * F5, *Qv = Qc + 3 * F5;
    double *Qw = Q;
    if (!c->p.mutate_q) {
        Qw = ws_get(c, 5);
        pcopy(Qw, Q, F5);
    }
    pcopy(Ux, Qw, F5);
    rhs_x(c, Ux, Qcx, Qv, W);
    rhs_yz(c, Qw, Ux, Qcy, Qcz, Qv, W);
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < F5; l++) Qrhs[l] = Qcx[l] + Qcy[l] + Qcz[l] + Qv[l];   /* :1381 */
    c->arena_top = mark;
}

void or_get_grad(const or_ctx *c, double *out12)
{
    memcpy(out12, c->grad, sizeof(double) * 12 * c->npts);
}

void or_get_hyper(const or_ctx *c, double *out3)
{
    memcpy(out3, c->hyp, sizeof(double) * 3 * c->npts);
}

void or_set_hyper(or_ctx *c, const double *in3)
{
    memcpy(c->hyp, in3, sizeof(double) * 3 * c->npts);
}

/* ---------------------------------------------------------------- */
/* artificial transport: src/hyper_dynamic.f                         */
/* ---------------------------------------------------------------- */
/* Compute_Laplacian :41-58: d2S = D2z + D2x + D2y (that association) */
static void laplacian(or_ctx *c, const double *S, double *d2S, double *tmp2)
{
    const long N = c->npts;
    double *t1 = tmp2, *t2 = tmp2 + N;
    or_d2fd2(c, 0, S, NULL, t1, 0);
    or_d2fd2(c, 1, S, NULL, t2, 0);
    or_d2fd2(c, 2, S, NULL, d2S, 0);
    PFOR d2S[l] = d2S[l] + t1[l] + t2[l];
}

/* Compute_coeff1 :72-165 (only C_beta varies; C_mu, C_k are constants) */
static void compute_coeff1(or_ctx *c, const double *Q, double *Cb)
{
    const long N = c->npts;
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3);
    const long mark = c->arena_top;
    double *U = ws_get(c, 3);
    double *G = ws_get(c, 4);
    double *tmp = ws_get(c, 1);
    PFOR {                                                   /* :87-97 */
        const double t = 1.0 / Q4[l];
        U[l] = Q1[l] * t;
        U[N + l] = Q2[l] * t;
        U[2 * N + l] = Q3[l] * t;
    }
    or_dfd(c, 1, U + 2 * N, G, 0, 1.0);                      /* :105 */
    or_dfd(c, 2, U + N, G, 1, -1.0);                         /* :106 */
    or_dfd(c, 0, U + 2 * N, G + N, 0, -1.0);                 /* :114 */
    or_dfd(c, 2, U, G + N, 1, 1.0);                          /* :115 */
    or_dfd(c, 0, U + N, G + 2 * N, 0, 1.0);                  /* :123 */
    or_dfd(c, 1, U, G + 2 * N, 1, -1.0);                     /* :124 */
    or_dfd(c, 0, U, G + 3 * N, 0, 1.0);                      /* :132-134 */
    or_dfd(c, 1, U + N, tmp, 0, 1.0);                        /* :136 */
    PFOR G[3 * N + l] = G[3 * N + l] + tmp[l];               /* :137 */
    or_dfd(c, 2, U + 2 * N, tmp, 0, 1.0);                    /* :139 */
    PFOR G[3 * N + l] = G[3 * N + l] + tmp[l];               /* :140 */
    const double eps = 1.0e-7, fac2 = 10.0;
    PFOR {                                                   /* :145-160 */
        const double g1 = G[l], g2 = G[N + l], g3 = G[2 * N + l], g4 = G[3 * N + l];
        const double fac = 5.0e-1 * (1.0 - tanh(2.5 + fac2 * g4));
        const double omega2 = g1 * g1 + g2 * g2 + g3 * g3;
        Cb[l] = fac * g4 * g4 / (g4 * g4 + omega2 + eps) + 1.0e-20;
    }
    c->arena_top = mark;
}

/* Compute_hyper :173-318 (uses /GRAD/ from the last or_rhs) */
void or_compute_hyper(or_ctx *c, const double *Q)
{
    const long N = c->npts;
    const double gamma = c->p.gamma;
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    const double *G = c->grad;
    const long mark = c->arena_top;
    double *Cb = ws_get(c, 1);
    double *P = ws_get(c, 1), *Vs = ws_get(c, 1);
    double *T = ws_get(c, 1), *e = ws_get(c, 1);
    double *S = ws_get(c, 1), *d2S = ws_get(c, 1);
    double *d4S = ws_get(c, 1), *d4e = ws_get(c, 1);
    double *tmp2 = ws_get(c, 2);
    double *eta = ws_get(c, 3), *etag = ws_get(c, 3);
    const double C_mu4 = 2.0e-3, C_k4 = 1.0e-2;              /* :162-163 */
    /* The non-periodic-x link set uses src/hyper_routines.f:66-214, the
       same chain with constant coefficients (C_beta4 = 1, :99-101) and two
       smoothed fields (:190-198). */
    const int stat = c->p.nonper_x;

    if (!stat) compute_coeff1(c, Q, Cb);                     /* :214 */
    /* :220-222 U = Q/rho is not used afterwards; omitted. */
    /* Sij (:227-245) and contraction (:269, Compute_Contraction :12-33) */
    PFOR {
        const double S11 = G[l], S22 = G[4 * N + l], S33 = G[8 * N + l];
        const double S12 = 5.0e-1 * (G[3 * N + l] + G[1 * N + l]);
        const double S13 = 5.0e-1 * (G[6 * N + l] + G[2 * N + l]);
        const double S23 = 5.0e-1 * (G[7 * N + l] + G[5 * N + l]);
        /* F2=S12 F4=S12, F3=S13 F7=S13, F6=S23 F8=S23 */
        S[l] = sqrt(5.0e-1 * (S11 * S11 + S22 * S22 + S33 * S33
                              + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    }
    const double fac1 = gamma - 1.0, fac2 = gamma / fac1;
    PFOR {                                                   /* :251-266 */
        Vs[l] = 1.0 / Q4[l];
        P[l] = fac1 * (Q5[l] - 5.0e-1 * Vs[l] * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
        T[l] = fac2 * Vs[l] * P[l];
        e[l] = Vs[l] * (Q5[l] - 5.0e-1 * Vs[l] * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
    }
    laplacian(c, S, d2S, tmp2);                              /* :272 */
    laplacian(c, d2S, d4S, tmp2);                            /* :273 */
    laplacian(c, e, d2S, tmp2);                              /* :275 */
    laplacian(c, d2S, d4e, tmp2);                            /* :276 */
    PFOR {                                                   /* :278-297 */
        const double a4S = fabs(d4S[l]), a4e = fabs(d4e[l]);
        const double fac3 = (Q4[l] / T[l]) * sqrt(gamma * P[l] * Vs[l]);
        eta[l] = Q4[l] * a4S;
        eta[2 * N + l] = Q4[l] * a4S;
        eta[N + l] = fac3 * a4e;
    }
    or_gaussian_filter(c, eta, etag, stat ? 2 : 3);          /* :300; static :198 */
    const double DX = c->D[0];
    double *beta = c->hyp, *visc = c->hyp + N, *kh = c->hyp + 2 * N;
    PFOR {                                                   /* :303-314; static :204-206 */
        beta[l] = stat ? 1.0 * etag[l] * pow(DX, 6.0) : Cb[l] * etag[2 * N + l] * pow(DX, 6.0);
        visc[l] = C_mu4 * etag[l] * pow(DX, 6.0);
        kh[l] = C_k4 * etag[N + l] * pow(DX, 5.0);
    }
    c->arena_top = mark;
}

/* Hyper_RHS src/hyper_rhs.f:12-499 */
void or_hyper_rhs(or_ctx *c, const double *Q, double *Qrhs)
{
    if (c->p.hyper_param == 0) return;                       /* :50-52 */
    const long N = c->npts;
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3);
    const double *G = c->grad;
    const double *dux_dx = G, *duy_dx = G + N, *duz_dx = G + 2 * N;
    const double *dux_dy = G + 3 * N, *duy_dy = G + 4 * N, *duz_dy = G + 5 * N;
    const double *dux_dz = G + 6 * N, *duy_dz = G + 7 * N, *duz_dz = G + 8 * N;
    const double *dTdx = G + 9 * N, *dTdy = G + 10 * N, *dTdz = G + 11 * N;
    or_compute_hyper(c, Q);                                  /* :57 */
    const double *beta = c->hyp, *visc = c->hyp + N, *kh = c->hyp + 2 * N;
    const long mark = c->arena_top;
    double *Qh = ws_get(c, 5);                                /* :54 */
    pset(Qh, 0.0, (size_t)5 * N);
    double *U = ws_get(c, 3);
    double *dila = ws_get(c, 1), *S = ws_get(c, 3);
    double *Flux = ws_get(c, 1), *FF = ws_get(c, 1);
    double *dFx = ws_get(c, 1), *dFy = ws_get(c, 1), *dFz = ws_get(c, 1);
    double *dEx = ws_get(c, 1), *dEy = ws_get(c, 1), *dEz = ws_get(c, 1);
    PFOR {                                                   /* :67-75 */
        U[l] = Q1[l] / Q4[l];
        U[N + l] = Q2[l] / Q4[l];
        U[2 * N + l] = Q3[l] / Q4[l];
    }
    PFOR dila[l] = dux_dx[l] + duy_dy[l] + duz_dz[l];        /* :83-89 */
    PFOR S[l] = dux_dy[l] + duy_dx[l];                       /* :95-101 */
    PFOR S[N + l] = dux_dz[l] + duz_dx[l];                   /* :107-113 */
    PFOR S[2 * N + l] = duy_dz[l] + duz_dy[l];               /* :119-125 */
    /* :146-157 Vs, P, T are not used afterwards; omitted. */
    const double tt = 2.0 / 3.0;
    /* x-momentum :175-253 */
    PFOR { Flux[l] = 2.0 * visc[l] * dux_dx[l] + (beta[l] - tt * visc[l]) * dila[l]; FF[l] = U[l] * Flux[l]; }
    or_dfd(c, 0, Flux, dFx, 0, 1.0); or_dfd(c, 0, FF, dEx, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[l]; FF[l] = U[l] * Flux[l]; }
    or_dfd(c, 1, Flux, dFy, 0, 1.0); or_dfd(c, 1, FF, dEy, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[N + l]; FF[l] = U[l] * Flux[l]; }
    or_dfd(c, 2, Flux, dFz, 0, 1.0); or_dfd(c, 2, FF, dEz, 0, 1.0);
    PFOR {
        Qh[l] = Qh[l] + dFx[l] + dFy[l] + dFz[l];
        Qh[4 * N + l] = Qh[4 * N + l] + dEx[l] + dEy[l] + dEz[l];
    }
    /* y-momentum :271-350 */
    PFOR { Flux[l] = 2.0 * visc[l] * duy_dy[l] + (beta[l] - tt * visc[l]) * dila[l]; FF[l] = U[N + l] * Flux[l]; }
    or_dfd(c, 1, Flux, dFy, 0, 1.0); or_dfd(c, 1, FF, dEy, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[l]; FF[l] = U[N + l] * Flux[l]; }
    or_dfd(c, 0, Flux, dFx, 0, 1.0); or_dfd(c, 0, FF, dEx, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[2 * N + l]; FF[l] = U[N + l] * Flux[l]; }
    or_dfd(c, 2, Flux, dFz, 0, 1.0); or_dfd(c, 2, FF, dEz, 0, 1.0);
    PFOR {
        Qh[N + l] = Qh[N + l] + dFx[l] + dFy[l] + dFz[l];
        Qh[4 * N + l] = Qh[4 * N + l] + dEx[l] + dEy[l] + dEz[l];
    }
    /* z-momentum :368-446 */
    PFOR { Flux[l] = 2.0 * visc[l] * duz_dz[l] + (beta[l] - tt * visc[l]) * dila[l]; FF[l] = U[2 * N + l] * Flux[l]; }
    or_dfd(c, 2, Flux, dFz, 0, 1.0); or_dfd(c, 2, FF, dEz, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[N + l]; FF[l] = U[2 * N + l] * Flux[l]; }
    or_dfd(c, 0, Flux, dFx, 0, 1.0); or_dfd(c, 0, FF, dEx, 0, 1.0);
    PFOR { Flux[l] = visc[l] * S[2 * N + l]; FF[l] = U[2 * N + l] * Flux[l]; }
    or_dfd(c, 1, Flux, dFy, 0, 1.0); or_dfd(c, 1, FF, dEy, 0, 1.0);
    PFOR {
        Qh[2 * N + l] = Qh[2 * N + l] + dFx[l] + dFy[l] + dFz[l];
        Qh[4 * N + l] = Qh[4 * N + l] + dEx[l] + dEy[l] + dEz[l];
    }
    /* heat conduction :454-491 */
    PFOR Flux[l] = kh[l] * dTdx[l];
    or_dfd(c, 0, Flux, dFx, 0, 1.0);
    PFOR Flux[l] = kh[l] * dTdy[l];
    or_dfd(c, 1, Flux, dFy, 0, 1.0);
    PFOR Flux[l] = kh[l] * dTdz[l];
    or_dfd(c, 2, Flux, dFz, 0, 1.0);
    PFOR Qh[4 * N + l] = Qh[4 * N + l] + dFx[l] + dFy[l] + dFz[l];
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < (size_t)5 * N; l++) Qrhs[l] = Qrhs[l] + Qh[l];   /* :494 */
    c->arena_top = mark;
}

/* ---------------------------------------------------------------- */
/* time integration: src/lddrk4_periodic_sf.f                        */
/* ---------------------------------------------------------------- */
/* Fix_Pressure :436-473 */
void or_fix_pressure(or_ctx *c, double *Q)
{
    const long N = c->npts;
    const double fac1 = c->p.gamma - 1.0;
    const double base_pressure = 0.71463;
    double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    long nclamp = 0;
#pragma omp parallel for schedule(static) reduction(+ : nclamp)
    for (long l = 0; l < N; l++) {
        const double Vs = 1.0 / Q4[l];
        double P = fac1 * (Q5[l] - 5.0e-1 * Vs * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
        if (P < 0.0) {
            P = base_pressure;
            nclamp++;
        }
        Q5[l] = (1.0 / fac1) * P + 5.0e-1 * Vs * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]);
    }
    c->clamps += nclamp;
}

long or_clamps(const or_ctx *c) { return c->clamps; }

/* Compute_CFL :7-65 */
void or_compute_cfl(or_ctx *c, const double *Q, double *stab, double *resv)
{
    const long N = c->npts;
    const double gamma = c->p.gamma;
    const double fac = gamma * (gamma - 1.0);
    const double DX = c->D[0], DY = c->D[1], DZ = c->D[2];
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    double m1 = -INFINITY, m2 = -INFINITY;
#pragma omp parallel for schedule(static) reduction(max : m1, m2)
    for (long l = 0; l < N; l++) {
        const double tmp1 = 1.0 / Q4[l];
        const double tmp2 = sqrt(fac * tmp1 * (Q5[l] - 5.0e-1 * tmp1 * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l])));
        const double ux = fabs(Q1[l] * tmp1), uy = fabs(Q2[l] * tmp1), uz = fabs(Q3[l] * tmp1);
        const double w1 = (tmp2 + ux) * k_max / DX + (tmp2 + uy) * k_max / DY + (tmp2 + uz) * k_max / DZ;
        const double w2 = (tmp2 + ux) * k_hug / DX + (tmp2 + uy) * k_hug / DY + (tmp2 + uz) * k_hug / DZ;
        if (w1 > m1) m1 = w1;
        if (w2 > m2) m2 = w2;
    }
    *stab = m1;
    *resv = m2;
}

/* Compute_Tstep :74-145.  The 0.2 safety factor is a default-real
   (single precision) literal in the reference; its double value is
   (double)0.2f. */
double or_compute_tstep(or_ctx *c, const double *Q, double cfl_resv, int step)
{
    const long N = c->npts;
    double Tstep = c->p.cfl_cal / cfl_resv;
    /* src/lddrk4_nonperiodic_sf.f:105 drops the step > 1 guard; before the
       first Compute_hyper the coefficients are zero and the limits infinite */
    if (c->p.hyper_param == 1 && (step > 1 || c->p.nonper_x)) {
        const double gamma = c->p.gamma;
        const double del = c->D[0];
        const double fac = gamma * (gamma - 1.0);
        const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
        const double *beta = c->hyp, *visc = c->hyp + N, *kh = c->hyp + 2 * N;
        double t3 = INFINITY, t4 = INFINITY, t5 = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : t3, t4, t5)
        for (long l = 0; l < N; l++) {
            const double tmp1 = 1.0 / Q4[l];
            const double tmp2 = sqrt(fac * tmp1 * (Q5[l] - 5.0e-1 * tmp1 * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l])));
            const double T = gamma * tmp1 * (Q5[l] - 5.0e-1 * tmp1 * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
            const double w3 = Q4[l] * pow(del, 2.0) / visc[l];
            const double w4 = Q4[l] * pow(del, 2.0) / beta[l];
            const double w5 = Q4[l] * pow(del * tmp2, 2.0) / (kh[l] * T);
            if (w3 < t3) t3 = w3;
            if (w4 < t4) t4 = w4;
            if (w5 < t5) t5 = w5;
        }
        const double p2 = (double)0.2f;
        double m = c->p.cfl_cal / cfl_resv;
        if (p2 * t3 < m) m = p2 * t3;
        if (p2 * t4 < m) m = p2 * t4;
        if (p2 * t5 < m) m = p2 * t5;
        Tstep = m;
    }
    return Tstep;
}

/* one stage body :344-356: RHS, Hyper_RHS, 2N-storage update, filter, fix */
void or_set_sponge(or_ctx *c, double amp, double power, double frac, const double *ref_left, const double *ref_right)
{
    c->sp_amp = amp;
    c->sp_pow = power;
    c->sp_frac = frac;
    for (int l = 0; l < 5; l++) {
        c->sp_ref[l] = ref_left ? ref_left[l] : 0.0;
        c->sp_ref[5 + l] = ref_right ? ref_right[l] : 0.0;
    }
}

/* Sponge_Forcing src/sponge1d.f:29-95 (one rank: gI = I); the left layer
   relaxes to the first reference state, the right layer to the second */
static void sponge(or_ctx *c, const double *Q, double *Qrhs)
{
    const int Nx = c->n[0];
    const long plane = (long)c->n[1] * c->n[2];
    const long N = c->npts;
    const int I_left = (int)ceil((c->sp_frac * Nx) / 2.0);          /* :35 */
    const int I_right = Nx - I_left;                                 /* :36 */
    const int I_len = I_left;                                        /* :38 */
#pragma omp parallel for schedule(static)
    for (int I = 1; I <= Nx; I++) {
        double sigma = 0.0;
        const double *ref = c->sp_ref;
        if (I <= I_left) {                                           /* :65-68 */
            const float r = (float)(I_left - I + 1) / (float)I_len;
            sigma = c->sp_amp * pow((double)r, c->sp_pow);
        } else if (I >= I_right) {                                   /* :69-72 */
            const float r = (float)(I - I_right) / (float)I_len;
            sigma = c->sp_amp * pow((double)r, c->sp_pow);
            ref = c->sp_ref + 5;
        }
        for (long p = 0; p < plane; p++) {
            const long l = (long)(I - 1) * plane + p;
            for (int L = 0; L < 5; L++)                                 /* :85-95 */
                Qrhs[L * N + l] = Qrhs[L * N + l] - (sigma) * (Q[L * N + l] - ref[L]);
        }
    }
}

void or_rk_stage(or_ctx *c, double *Q, double *QW, double *dQ, double alpha, double beta, double dt)
{
    const size_t F5 = (size_t)5 * c->npts;
    or_rhs(c, Q, dQ);
    or_hyper_rhs(c, Q, dQ);
    if (c->sp_amp != 0.0) sponge(c, Q, dQ);       /* hook order of src/lddrk4_nonperiodic_mf.f:462-466 */
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < F5; l++) {
        QW[l] = alpha * QW[l] + dt * dQ[l];                  /* :349 */
        Q[l] = Q[l] + beta * QW[l];                          /* :350 */
    }
    const long mark = c->arena_top;
    double *Qf = ws_get(c, 5);
    or_compact_filter(c, Q, Qf, 5);                          /* :353 */
    pcopy(Q, Qf, F5);                                        /* :354 */
    c->arena_top = mark;
    or_fix_pressure(c, Q);                                   /* :356 */
}

/* LDDRK4 :311-432: 5-stage loop then 6-stage loop (time += dt after each) */
void or_lddrk4(or_ctx *c, double *Q, double dt, double *time, int *step)
{
    const size_t F5 = (size_t)5 * c->npts;
    const long mark = c->arena_top;
    double *QW = ws_get(c, 5), *dQ = ws_get(c, 5);
    for (int loop = 0; loop < 2; loop++) {
        double a[6], b[6];
        const int ns = or_rk_coeffs(c->p.rk_scheme, loop, a, b);
        if (ns == 0) break;
        pset(QW, 0.0, F5);                                   /* :339, :376 */
        for (int k = 0; k < ns; k++) or_rk_stage(c, Q, QW, dQ, a[k], b[k], dt);
        *time = *time + dt;                                  /* :372, :419 */
    }
    *step = *step + 1;                                       /* :420 */
    c->arena_top = mark;
}

/* src/main.f:129-136 */
double or_march(or_ctx *c, double *Q, double *time, int *step)
{
    double stab, resv;
    or_compute_cfl(c, Q, &stab, &resv);
    const double dt = or_compute_tstep(c, Q, resv, *step);
    or_lddrk4(c, Q, dt, time, step);
    return dt;
}

/* Compute_Sum src/auxiliary.f:187-216: loop I (x) outer, J (y), K (z) inner */
static double compute_sum(const or_ctx *c, const double *F)
{
    const long My = c->n[1], Mz = c->n[2], Mx = c->n[0];
    double s = 0.0;
    for (long i = 0; i < Mx; i++)
        for (long j = 0; j < My; j++)
            for (long k = 0; k < Mz; k++) s = s + F[j + My * (k + Mz * i)];
    return s;
}

/* Compute_Mom/KE/Mass/Energy/Entropy :153-302 */
/* Compute_Grad src/grad.f:13-115: the nine velocity gradients of /GRAD/
   from Q, MaxDil = max |dil| (Compute_Max) and Enstrophy = 0.5 sum
   |omega|^2 / (Nx Ny Nz) (Compute_Sum, src/auxiliary.f:187-216) */
void or_compute_grad(or_ctx *c, const double *Q, double *maxdil, double *enstrophy)
{
    const long N = c->npts;
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3);
    double *G = c->grad;
    double *dux_dx = G, *duy_dx = G + N, *duz_dx = G + 2 * N;
    double *dux_dy = G + 3 * N, *duy_dy = G + 4 * N, *duz_dy = G + 5 * N;
    double *dux_dz = G + 6 * N, *duy_dz = G + 7 * N, *duz_dz = G + 8 * N;
    const long mark = c->arena_top;
    double *U = ws_get(c, 3), *Gr = ws_get(c, 4), *W = ws_get(c, 1);
    PFOR {                                                   /* :40-48 */
        const double tmp = 1.0 / Q4[l];
        U[l] = Q1[l] * tmp;
        U[N + l] = Q2[l] * tmp;
        U[2 * N + l] = Q3[l] * tmp;
    }
    or_dfd(c, 1, U + 2 * N, duz_dy, 0, 1.0);                 /* :55-57 */
    or_dfd(c, 2, U + N, duy_dz, 0, 1.0);
    PFOR Gr[l] = duz_dy[l] - duy_dz[l];
    or_dfd(c, 0, U + 2 * N, duz_dx, 0, 1.0);                 /* :66-68 */
    or_dfd(c, 2, U, dux_dz, 0, 1.0);
    PFOR Gr[N + l] = dux_dz[l] - duz_dx[l];
    or_dfd(c, 0, U + N, duy_dx, 0, 1.0);                     /* :77-79 */
    or_dfd(c, 1, U, dux_dy, 0, 1.0);
    PFOR Gr[2 * N + l] = duy_dx[l] - dux_dy[l];
    or_dfd(c, 0, U, dux_dx, 0, 1.0);                         /* :88-96 */
    or_dfd(c, 1, U + N, duy_dy, 0, 1.0);
    or_dfd(c, 2, U + 2 * N, duz_dz, 0, 1.0);
    PFOR Gr[3 * N + l] = dux_dx[l] + duy_dy[l] + duz_dz[l];
    double m = 0.0;                                          /* :100-104 */
    for (long l = 0; l < N; l++) {
        const double a = fabs(Gr[3 * N + l]);
        if (a > m) m = a;
    }
    PFOR W[l] = Gr[l] * Gr[l] + Gr[N + l] * Gr[N + l] + Gr[2 * N + l] * Gr[2 * N + l];   /* :110-111 */
    const double NN = (double)(c->n[0] * c->n[1] * c->n[2]);
    if (maxdil) *maxdil = m;
    if (enstrophy) *enstrophy = 0.5 * compute_sum(c, W) / NN;   /* :113-115 */
    c->arena_top = mark;
}

void or_diagnostics(or_ctx *c, const double *Q, double *out)
{
    const long N = c->npts;
    const double *Q1 = FLD(Q, 0), *Q2 = FLD(Q, 1), *Q3 = FLD(Q, 2), *Q4 = FLD(Q, 3), *Q5 = FLD(Q, 4);
    const double NN = (double)(c->n[0] * c->n[1] * c->n[2]);
    const long mark = c->arena_top;
    double *tmp = ws_get(c, 1);
    out[0] = compute_sum(c, Q1) / NN;
    out[1] = compute_sum(c, Q2) / NN;
    out[2] = compute_sum(c, Q3) / NN;
    PFOR {
        const double t = 1.0 / Q4[l];
        const double u1 = Q1[l] * t, u2 = Q2[l] * t, u3 = Q3[l] * t;
        double s = 0.0;
        s = s + Q4[l] * u1 * u1;
        s = s + Q4[l] * u2 * u2;
        s = s + Q4[l] * u3 * u3;
        tmp[l] = s;
    }
    out[3] = 0.5 * compute_sum(c, tmp) / NN;
    out[4] = compute_sum(c, Q4) / NN;
    out[5] = compute_sum(c, Q5) / NN;
    const double gamma = c->p.gamma, fac1 = gamma - 1.0;
    PFOR {
        const double Vs = 1.0 / Q4[l];
        const double P = fac1 * (Q5[l] - 5.0e-1 * Vs * (Q1[l] * Q1[l] + Q2[l] * Q2[l] + Q3[l] * Q3[l]));
        tmp[l] = Q4[l] * log(P * pow(Vs, gamma));
    }
    out[6] = compute_sum(c, tmp) / NN;
    c->arena_top = mark;
}
