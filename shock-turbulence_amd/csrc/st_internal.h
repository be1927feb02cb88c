// Internal declarations shared by the HIP kernels and the host side of the
// C-ABI.  Nothing here crosses the library boundary.
#pragma once

#include <cstddef>
#include <cstdint>

namespace st {

// Cyclic constant-band solve of (beta, alpha, 1, alpha, beta) as a product of
// a forward and a backward linear recurrence.  Order 1 (beta = 0):
//   A = (1/(1+rho^2)) (1 - rho S)(1 - rho S^-1),  S = cyclic shift,
// order 2: A = kappa (1 - a1 S - a2 S^2)(1 - a1 S^-1 - a2 S^-2).
// Each lane of a wave owns M consecutive points of a line; P lanes own a
// line of n = M*P points.  Chunk-local recurrences are stitched together by
// a segmented scan over the P lanes.  With the companion matrix
// A = [[a1, a2], [1, 0]] (order 1: a2 = 0) the tables hold:
struct OpTab {
    int order;            // 1 or 2
    int pad_;
    double a1, a2;        // recurrence coefficients
    double scale;         // folded into the right-hand side: 1+rho^2 or 1/kappa
    double cf[16][2];     // row 0 of A^(k+1), k = 0..M-1 (in-chunk carry weights)
    double sc[6][4];      // A^(M*2^j), j = 0..log2(P)-1 (scan steps, row-major 2x2)
    double wrap[4];       // (I - A^n)^-1
    double lp[64][4];     // A^(M*c), c = 0..P-1 (per-lane injection of the cyclic total)
    // Non-periodic lines (x of the shock-tube link set).  The first and
    // last rows of the system are boundary closures; the recurrences run
    // without the cyclic wrap and so invert Ah = kappa L U (L, U the
    // recurrence matrices truncated to the line).  The closure matrix is
    // A = Ah + sum_j e_(i_j) v_j^T over the np_m boundary rows i_j, and
    //   x = y - Z q,  q = G yb,  y = Ah^-1 b,
    // with Z = Ah^-1 [e_(i_j)] and G = (I + V^T Z)^-1 V^T restricted to the
    // ten boundary values yb = (y_0..y_4, y_(n-5)..y_(n-1)) (Woodbury); Z q
    // is one more solve of the boundary rows, so only G is tabled.
    int np_m;             // boundary rows (4: first/second derivative, 6: filter)
    int pad2_;
    double npG[6][10];
};

// Per-direction operator tables and stencil factors.
struct DirTab {
    OpTab op[3];          // 0: first derivative, 1: second derivative, 2: compact filter
    double s1;            // sx(1)*(N-1)        (src/derivCDopt6_periodic.f:181)
    double s2;            // (N-1)^2*sx(2)      (src/derivCDopt6_periodic.f:298)
    int n, M, P, B;       // line length, points per lane, lanes per line, lines per tile
    int nonper;           // 1: closure rows at both ends, no wrap (the tables' np_* fields)
    int pad_;
};

// Compact-scheme coefficients (src/header:191-203, src/LES_header:104-157).
struct Coef {
    double a1, b1, c1;            // D1 right-hand side
    double a2, b2, c2, d2;        // D2 right-hand side (+-1, +-2, +-3, centre)
    double f0, f1, f2, f3, f4;    // compact filter right-hand side
    double g0, g1, g2, g3, g4;    // Gaussian
};

// Physical parameters used in kernels.
struct Phys {
    double gamma, gm1, gfac, gm1_inv;   // gamma, gamma-1, gamma/(gamma-1), 1/(gamma-1)
    double ginv;                        // 1/gamma
    double re_inv, pr_inv, vis_exp, vis_exp_m1, bulk;
    double dx6, dx5;                    // DX**6, DX**5 (src/hyper_dynamic.f:307-311)
    double cmu, ck;                     // C_mu = 2e-3, C_k = 1e-2
    double D[3];                        // DX, DY, DZ
    double cfl_cal;
};

// Fill the three operator tables of one direction for a line of n points
// split into P lanes of M points.  Returns 0 on success.
// nonper = 1 adds the boundary-closure tables of the non-periodic x
// operators (src/derivCDopt6_nonperiodic1d.f:41-357,
// src/filters_nonperiodic1d.f:45-215).
int build_dirtab(DirTab *t, int n, int M, int P, int B, double sx1, double sx2,
                 double d1_alpha, double d2_alpha, double f_alpha, double f_beta, int nonper = 0);

// Choose (M, P) for a line length; returns 0 if n is unsupported.
int choose_layout(int n, int *M, int *P);

// Setup_Grid of one direction (src/grid.f:31-40, src/define_periodic_sf.f:42-44):
// x_i = start + leng (i-1)/(n-1) with leng = period (n-1)/n, or the whole
// period with both end points for a non-periodic direction
// (src/define_nonperiodic1d_sf.f:42); the spacing is x_2 - x_1.
double grid_leng(double period, int n, int nonper);
double grid_spacing(double period, double start, int n, int nonper);

// Failure of a call that has no context: records the message (per thread)
// for st_last_error(NULL) and returns code.
int set_call_error(int code, const char *fmt, ...);

// Seeded isotropic turbulence initial condition (host).
int ic_isotropic(int nx, int ny, int nz, const double period[3], double k0, double urms,
                 double mt, double gamma, unsigned long long seed, double *Q);
// Planar Mach-M shock at x = frac L with upstream turbulence (host).
int ic_shock(int nx, int ny, int nz, const double period[3], double mach, double frac, double width,
             double k0, double urms, double gamma, unsigned long long seed, double *Q, double *states);

// Sponge strength of Sponge_Forcing (src/sponge1d.f:29-79) at x index
// i = 0..nx-1: amp ((d/len))^power over the first and last ceil(frac nx/2)
// points, the ratio formed in single precision as the reference's REAL().
double sponge_sigma(int nx, int i, double amp, double power, double frac);

}  // namespace st
