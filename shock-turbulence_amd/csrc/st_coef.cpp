// Host precomputation of the recurrence tables used by the line-solve
// kernels (see OpTab in st_internal.h).  All powers are formed in long
// double and rounded once.
#include <cmath>
#include <complex>
#include <cstring>
#include <utility>
#include <vector>

#include "st_internal.h"

namespace st {

namespace {

typedef long double LD;

struct M2 {
    LD a, b, c, d;   // [[a, b], [c, d]]
};

M2 mul(const M2 &x, const M2 &y)
{
    return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}

M2 mpow(const M2 &x, long e)
{
    M2 r = {1, 0, 0, 1}, b = x;
    while (e > 0) {
        if (e & 1) r = mul(r, b);
        b = mul(b, b);
        e >>= 1;
    }
    return r;
}

void put(double *dst, const M2 &m)
{
    dst[0] = (double)m.a;
    dst[1] = (double)m.b;
    dst[2] = (double)m.c;
    dst[3] = (double)m.d;
}

// Factor 1 + alpha (z + 1/z) + beta (z^2 + 1/z^2) = kappa (1 - a1 z - a2 z^2)(1 - a1/z - a2/z^2).
int factor(double alpha, double beta, LD *a1, LD *a2, LD *kappa)
{
    if (beta == 0.0) {
        if (alpha == 0.0) { *a1 = 0; *a2 = 0; *kappa = 1; return 0; }
        const LD al = alpha;
        const LD disc = 1.0L - 4.0L * al * al;
        if (disc <= 0) return -1;
        const LD rho = (-1.0L + std::sqrt(disc)) / (2.0L * al);
        *a1 = rho;
        *a2 = 0;
        *kappa = 1.0L / (1.0L + rho * rho);
        return 0;
    }
    typedef std::complex<LD> C;
    const LD al = alpha, be = beta;
    // s = z + 1/z solves be s^2 + al s + (1 - 2 be) = 0
    const C disc = C(al * al - 4.0L * be * (1.0L - 2.0L * be), 0.0L);
    const C sq = std::sqrt(disc);
    C s[2] = {(-al + sq) / (2.0L * be), (-al - sq) / (2.0L * be)};
    C z[2];
    for (int i = 0; i < 2; i++) {
        const C r = std::sqrt(s[i] * s[i] - 4.0L);
        C z1 = (s[i] + r) / 2.0L, z2 = (s[i] - r) / 2.0L;
        z[i] = std::abs(z1) < std::abs(z2) ? z1 : z2;
        if (!(std::abs(z[i]) < 1.0L)) return -2;
    }
    *a1 = std::real(z[0] + z[1]);
    *a2 = -std::real(z[0] * z[1]);
    *kappa = -be / *a2;
    const LD chk = *kappa * (1.0L + *a1 * *a1 + *a2 * *a2);
    if (std::fabs(chk - 1.0L) > 1e-12L) return -3;
    return 0;
}

int fill_op(OpTab *t, double alpha, double beta, int n, int M, int P)
{
    LD a1, a2, kappa;
    if (factor(alpha, beta, &a1, &a2, &kappa)) return -1;
    std::memset(t, 0, sizeof(*t));
    t->order = beta == 0.0 ? 1 : 2;
    t->a1 = (double)a1;
    t->a2 = (double)a2;
    t->scale = (double)(1.0L / kappa);
    const M2 A = {a1, a2, 1, 0};
    for (int k = 0; k < M; k++) {
        const M2 p = mpow(A, k + 1);
        t->cf[k][0] = (double)p.a;
        t->cf[k][1] = (double)p.b;
    }
    int j = 0;
    for (int d = 1; d < P; d *= 2, j++) put(t->sc[j], mpow(A, (long)M * d));
    for (int c = 0; c < P; c++) put(t->lp[c], mpow(A, (long)M * c));
    const M2 An = mpow(A, n);
    const M2 I_An = {1 - An.a, -An.b, -An.c, 1 - An.d};
    const LD det = I_An.a * I_An.d - I_An.b * I_An.c;
    const M2 inv = {I_An.d / det, -I_An.b / det, -I_An.c / det, I_An.a / det};
    put(t->wrap, inv);
    return 0;
}

// ---- non-periodic lines ---------------------------------------------

// LHS band row i (lower-2, lower-1, diagonal, upper+1, upper+2) of the
// non-periodic operator op: 0 first derivative
// (src/derivCDopt6_nonperiodic1d.f:70-158), 1 second derivative (:232-354),
// 2 compact filter (src/filters_nonperiodic1d.f:73-174, alpha_b = 0.45 from
// src/LES_header:118).  Interior rows are the periodic band.
void np_row(int op, int n, int i, double alpha, double beta, LD r[5])
{
    r[0] = beta; r[1] = alpha; r[2] = 1; r[3] = alpha; r[4] = beta;
    // mirror the last rows onto the first
    const bool end = i >= n - 4;
    const int j = end ? n - 1 - i : i;
    LD lo = -1, hi = -1, b2 = -1;      // -1: keep the interior row
    if (op == 0) {
        if (j == 0) { lo = 0; hi = 2; }
        else if (j == 1) { lo = 1.0L / 6.0L; hi = 0.5L; }
    } else if (op == 1) {
        if (j == 0) { lo = 0; hi = 11; }
        else if (j == 1) { lo = 0.1L; hi = -7.0L / 20.0L; }
    } else {
        const LD ab = 0.45L;
        if (j == 0) { lo = 0; hi = 0; b2 = 0; }
        else if (j == 1 || j == 2) { lo = ab; hi = ab; b2 = 0; }
    }
    if (lo < -0.5L) return;
    if (b2 > -0.5L) { r[0] = b2; r[4] = b2; }
    r[1] = end ? hi : lo;
    r[3] = end ? lo : hi;
}

// Solve the m x m system (I + VZ) G = V in place (Gauss-Jordan, partial
// pivoting); a holds the m x m matrix, g the m x 10 right-hand sides.
int gj_solve(int m, LD a[6][6], LD g[6][10])
{
    for (int col = 0; col < m; col++) {
        int piv = col;
        for (int r = col + 1; r < m; r++)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (!(std::fabs(a[piv][col]) > 1e-30L)) return -1;
        if (piv != col) {
            for (int k = 0; k < m; k++) std::swap(a[piv][k], a[col][k]);
            for (int k = 0; k < 10; k++) std::swap(g[piv][k], g[col][k]);
        }
        const LD inv = 1.0L / a[col][col];
        for (int k = 0; k < m; k++) a[col][k] *= inv;
        for (int k = 0; k < 10; k++) g[col][k] *= inv;
        for (int r = 0; r < m; r++) {
            if (r == col) continue;
            const LD f = a[r][col];
            if (f == 0) continue;
            for (int k = 0; k < m; k++) a[r][k] -= f * a[col][k];
            for (int k = 0; k < 10; k++) g[r][k] -= f * g[col][k];
        }
    }
    return 0;
}

// Boundary-correction table G of one operator (the np_* fields of OpTab),
// built from the rounded recurrence coefficients the kernels use (the
// columns of Z only on the host: the kernels apply Z through the same
// recurrences).
int fill_np(OpTab *t, int op, double alpha, double beta, int n)
{
    if (n < 16 || n > 512) return -1;
    const int m = op == 2 ? 6 : 4;
    int rows[6];
    for (int j = 0; j < m / 2; j++) {
        rows[j] = j;
        rows[m / 2 + j] = n - m / 2 + j;
    }
    const LD a1 = t->a1, a2 = t->a2, sc = t->scale;
    // row i of Ah = L U / scale with L = I - a1 S - a2 S^2 (S the down shift)
    // and U its transpose, both truncated to the line
    auto ah = [&](int i, int c) -> LD {
        LD s = 0;
        for (int k = i - 2; k <= i; k++) {
            if (k < 0 || k >= n) continue;
            const LD l = k == i ? 1 : (k == i - 1 ? -a1 : -a2);
            const int d = c - k;
            if (d < 0 || d > 2) continue;
            const LD u = d == 0 ? 1 : (d == 1 ? -a1 : -a2);
            s += l * u;
        }
        return s / sc;
    };
    // V rows: E = A - Ah on the boundary rows (columns i-2..i+2)
    LD V[6][5];
    for (int j = 0; j < m; j++) {
        LD r[5];
        np_row(op, n, rows[j], alpha, beta, r);
        for (int q = 0; q < 5; q++) {
            const int c = rows[j] - 2 + q;
            V[j][q] = (c < 0 || c >= n) ? 0 : r[q] - ah(rows[j], c);
        }
    }
    // Z columns: scale * U^-1 L^-1 e_i by the kernels' recurrences
    std::vector<LD> Z((size_t)m * n);
    for (int j = 0; j < m; j++) {
        LD *z = &Z[(size_t)j * n];
        for (int k = 0; k < n; k++) z[k] = k == rows[j] ? sc : 0;
        for (int k = 1; k < n; k++) z[k] += a1 * z[k - 1] + (k >= 2 ? a2 * z[k - 2] : 0);
        for (int k = n - 2; k >= 0; k--) z[k] += a1 * z[k + 1] + (k <= n - 3 ? a2 * z[k + 2] : 0);
    }
    LD IVZ[6][6], G[6][10];
    for (int j = 0; j < m; j++) {
        for (int l = 0; l < m; l++) {
            LD s = j == l ? 1 : 0;
            for (int q = 0; q < 5; q++) {
                const int c = rows[j] - 2 + q;
                if (c >= 0 && c < n) s += V[j][q] * Z[(size_t)l * n + c];
            }
            IVZ[j][l] = s;
        }
        for (int tt = 0; tt < 10; tt++) {
            const int c = tt < 5 ? tt : n - 10 + tt;
            const int q = c - rows[j] + 2;
            G[j][tt] = (q >= 0 && q < 5) ? V[j][q] : 0;
        }
    }
    if (gj_solve(m, IVZ, G)) return -2;
    t->np_m = m;
    for (int j = 0; j < m; j++)
        for (int tt = 0; tt < 10; tt++) t->npG[j][tt] = (double)G[j][tt];
    return 0;
}

}  // namespace

double grid_leng(double period, int n, int nonper)
{
    return nonper ? period : period * (double)(n - 1) / (double)n;
}

double grid_spacing(double period, double start, int n, int nonper)
{
    const double leng = grid_leng(period, n, nonper);
    const double g1 = start + leng * 0.0 / (double)(n - 1);
    const double g2 = start + leng * 1.0 / (double)(n - 1);
    return g2 - g1;
}

double sponge_sigma(int nx, int i, double amp, double power, double frac)
{
    // src/sponge1d.f:33-36: CEILING((frac_sponge_x*Nx)/2D0) points at each end
    const int start_left = (int)std::ceil((frac * nx) / 2.0);
    const int start_right = nx - start_left;
    const int len = start_left;
    const int g = i + 1;                                   // global 1-based index (:63)
    if (len < 1) return 0.0;
    // :65-73: ( REAL(a)/REAL(len) )**s_exp, the quotient in single precision
    if (g <= start_left) return amp * std::pow((double)((float)(start_left - g + 1) / (float)len), power);
    if (g >= start_right) return amp * std::pow((double)((float)(g - start_right) / (float)len), power);
    return 0.0;
}

int choose_layout(int n, int *M, int *P)
{
    static const int cand[] = {8, 7, 6, 5, 4, 2};
    for (int m : cand) {
        if (n % m) continue;
        const int p = n / m;
        if (p == 8 || p == 16 || p == 32 || p == 64) {
            *M = m;
            *P = p;
            return 1;
        }
    }
    return 0;
}

int build_dirtab(DirTab *t, int n, int M, int P, int B, double sx1, double sx2,
                 double d1_alpha, double d2_alpha, double f_alpha, double f_beta, int nonper)
{
    if (fill_op(&t->op[0], d1_alpha, 0.0, n, M, P)) return -1;
    if (fill_op(&t->op[1], d2_alpha, 0.0, n, M, P)) return -2;
    if (fill_op(&t->op[2], f_alpha, f_beta, n, M, P)) return -3;
    t->nonper = nonper ? 1 : 0;
    if (nonper) {
        if (fill_np(&t->op[0], 0, d1_alpha, 0.0, n)) return -4;
        if (fill_np(&t->op[1], 1, d2_alpha, 0.0, n)) return -5;
        if (fill_np(&t->op[2], 2, f_alpha, f_beta, n)) return -6;
    }
    t->s1 = sx1 * (double)(n - 1);
    t->s2 = (double)((n - 1) * (n - 1)) * sx2;
    t->n = n;
    t->M = M;
    t->P = P;
    t->B = B;
    return 0;
}

}  // namespace st
