// Host precomputation of the recurrence tables used by the line-solve
// kernels (see OpTab in st_internal.h).  All powers are formed in long
// double and rounded once.
#include <cmath>
#include <complex>
#include <cstring>

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

}  // namespace

int choose_layout(int n, int *M, int *P)
{
    static const int cand[] = {8, 4, 2};
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
                 double d1_alpha, double d2_alpha, double f_alpha, double f_beta)
{
    if (fill_op(&t->op[0], d1_alpha, 0.0, n, M, P)) return -1;
    if (fill_op(&t->op[1], d2_alpha, 0.0, n, M, P)) return -2;
    if (fill_op(&t->op[2], f_alpha, f_beta, n, M, P)) return -3;
    t->s1 = sx1 * (double)(n - 1);
    t->s2 = (double)((n - 1) * (n - 1)) * sx2;
    t->n = n;
    t->M = M;
    t->P = P;
    t->B = B;
    return 0;
}

}  // namespace st
