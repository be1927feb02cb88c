// Line-pass engine for CDNA4 (gfx950), FP64.
//
// A pass sweeps the whole grid along one direction.  A workgroup of NT
// threads owns a tile of B complete lines (B*n = NT*M points):
//   1. load:    the tile's input fields are read with coalesced global
//               loads (memory order) and staged in LDS;
//   2. compute: each wave holds 64/P lines, each lane M consecutive points
//               of one line; stencils take their halo from neighbouring
//               lanes by shuffles, and cyclic band solves run as
//               chunk-local recurrences stitched by segmented wave scans
//               (no barriers inside a solve);
//   3. store:   results go back through LDS and leave with coalesced
//               stores, optionally combined with other fields at the point.
#pragma once

#include <hip/hip_runtime.h>

#include "st_internal.h"

#define DEV __device__ __forceinline__
#define HD __host__ __device__ __forceinline__

namespace st {

constexpr int HALO = 4;

// Tile geometry of one pass launch.
struct Geo {
    int dir;               // 0: x (stride My*Mz), 1: y (stride 1), 2: z (stride My)
    int n, P, B;           // line length, lanes per line, lines per tile
    int My, Mz, Mx;
    long ntiles;
    long npts;
};

// Global index of element s of line `line` in direction dir.
DEV long gindex(const Geo &g, long line, int s)
{
    if (g.dir == 0) return line + (long)s * ((long)g.My * g.Mz);
    if (g.dir == 1) return line * g.My + s;
    const long j = line % g.My, i = line / g.My;
    return j + (long)g.My * g.Mz * i + (long)s * g.My;
}

// Tile for this workgroup.  Workgroups are dispatched round-robin over the
// 8 XCDs; consecutive tiles (which share 128-byte lines for the strided
// directions) are handed to workgroups of the same XCD.
DEV long tile_of(const Geo &g)
{
    const long wg = blockIdx.x;
    if ((g.ntiles & 7) == 0) return (wg & 7) * (g.ntiles >> 3) + (wg >> 3);
    return wg;
}

// LDS layout of one slot: line b holds its n = M*P points in a block of
// n doubles; point s = c*M + k sits at row k, column c.  Rows are rotated by
// b and columns XOR-swizzled so that lane-parallel reads at fixed k and the
// memory-order staging writes both spread over the 32 double-wide banks.
template <int M>
HD int lds_idx(int b, int k, int c, int P)
{
    const int f = k * (P / M) + (M == 4 ? 4 * b : b);
    return b * (M * P) + (((k + b) * P + (c ^ (f & (P - 1)))) & (M * P - 1));
}

template <int M>
HD int slot_size(int P, int B)
{
    return B * M * P;
}

// Element e (0 .. B*n-1) of the tile in memory order -> (b, s).
DEV void tile_elem(const Geo &g, int e, int &b, int &s)
{
    if (g.dir == 1) {
        b = e / g.n;
        s = e - b * g.n;
    } else {
        b = e & (g.B - 1);
        s = e / g.B;
    }
}

// Per-lane view during the compute phase.
template <int M>
struct Lane {
    int c, P, seg, b;      // chunk index, lanes per line, first lane of the line, line in tile
    double *lds;
    int ss;                // slot size (doubles)

    DEV double ld(int slot, int k) const { return lds[slot * ss + lds_idx<M>(b, k, c, P)]; }
    DEV void st(int slot, int k, double v) const { lds[slot * ss + lds_idx<M>(b, k, c, P)] = v; }
    DEV int src(int dc) const { return seg + ((c + dc) & (P - 1)); }
};

template <int M>
DEV Lane<M> make_lane(const Geo &g, double *lds)
{
    Lane<M> L;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    L.P = g.P;
    L.c = lane & (g.P - 1);
    L.seg = lane - L.c;
    L.b = wave * (64 / g.P) + lane / g.P;
    L.lds = lds;
    L.ss = slot_size<M>(g.P, g.B);
    return L;
}

// Window of M own values plus HALO on each side, fetched from neighbours.
template <int M>
DEV void window(const double (&v)[M], const Lane<M> &L, double (&w)[M + 2 * HALO])
{
#pragma unroll
    for (int k = 0; k < M; k++) w[HALO + k] = v[k];
#pragma unroll
    for (int o = 1; o <= HALO; o++) {
        const int dl = (o + M - 1) / M;
        const int idx = dl * M - o;
        w[HALO - o] = __shfl(v[idx], L.src(-dl));
    }
#pragma unroll
    for (int o = 0; o < HALO; o++) {
        const int dl = 1 + o / M;
        const int idx = o % M;
        w[HALO + M + o] = __shfl(v[idx], L.src(dl));
    }
}

// ---- cyclic solves ------------------------------------------------------

// Order 1: x = (1+rho^2) (1 - rho S^-1)^-1 (1 - rho S)^-1 r; the factor
// (1+rho^2) is already folded into r.
template <int M>
DEV void solve1(double (&r)[M], const OpTab &T, const Lane<M> &L)
{
    const double rho = T.a1;
    const int P = L.P, c = L.c;
#pragma unroll
    for (int k = 1; k < M; k++) r[k] = fma(rho, r[k - 1], r[k]);
    double E = r[M - 1];
    {
        int j = 0;
        for (int d = 1; d < P; d <<= 1, j++) {
            const double t = __shfl(E, L.src(-d));
            if (c >= d) E = fma(T.sc[j][0], t, E);
        }
    }
    const double Etot = __shfl(E, L.seg + P - 1) * T.wrap[0];
    double Ep = __shfl(E, L.src(-1));
    if (c == 0) Ep = 0.0;
    const double cin = fma(T.lp[c][0], Etot, Ep);
#pragma unroll
    for (int k = 0; k < M; k++) r[k] = fma(T.cf[k][0], cin, r[k]);
#pragma unroll
    for (int k = M - 2; k >= 0; k--) r[k] = fma(rho, r[k + 1], r[k]);
    double Bv = r[0];
    {
        int j = 0;
        for (int d = 1; d < P; d <<= 1, j++) {
            const double t = __shfl(Bv, L.src(d));
            if (c + d < P) Bv = fma(T.sc[j][0], t, Bv);
        }
    }
    const double Btot = __shfl(Bv, L.seg) * T.wrap[0];
    double Bn = __shfl(Bv, L.src(1));
    if (c == P - 1) Bn = 0.0;
    const double bin = fma(T.lp[P - 1 - c][0], Btot, Bn);
#pragma unroll
    for (int k = 0; k < M; k++) r[k] = fma(T.cf[M - 1 - k][0], bin, r[k]);
}

DEV void mv(const double *m, double x0, double x1, double &y0, double &y1)
{
    y0 = fma(m[0], x0, m[1] * x1);
    y1 = fma(m[2], x0, m[3] * x1);
}

// Order 2: x = (1/kappa) (1 - a1 S^-1 - a2 S^-2)^-1 (1 - a1 S - a2 S^2)^-1 r.
template <int M>
DEV void solve2(double (&r)[M], const OpTab &T, const Lane<M> &L)
{
    const double a1 = T.a1, a2 = T.a2;
    const int P = L.P, c = L.c;
    if (M > 1) r[1] = fma(a1, r[0], r[1]);
#pragma unroll
    for (int k = 2; k < M; k++) r[k] = fma(a1, r[k - 1], fma(a2, r[k - 2], r[k]));
    double e0 = r[M - 1], e1 = M > 1 ? r[M > 1 ? M - 2 : 0] : 0.0;
    {
        int j = 0;
        for (int d = 1; d < P; d <<= 1, j++) {
            const double t0 = __shfl(e0, L.src(-d)), t1 = __shfl(e1, L.src(-d));
            if (c >= d) {
                double y0, y1;
                mv(T.sc[j], t0, t1, y0, y1);
                e0 += y0;
                e1 += y1;
            }
        }
    }
    double w0 = __shfl(e0, L.seg + P - 1), w1 = __shfl(e1, L.seg + P - 1);
    double tot0, tot1;
    mv(T.wrap, w0, w1, tot0, tot1);
    double p0 = __shfl(e0, L.src(-1)), p1 = __shfl(e1, L.src(-1));
    if (c == 0) { p0 = 0.0; p1 = 0.0; }
    double q0, q1;
    mv(T.lp[c], tot0, tot1, q0, q1);
    const double cin0 = p0 + q0, cin1 = p1 + q1;
#pragma unroll
    for (int k = 0; k < M; k++) r[k] = fma(T.cf[k][0], cin0, fma(T.cf[k][1], cin1, r[k]));
    if (M > 1) r[M - 2] = fma(a1, r[M - 1], r[M - 2]);
#pragma unroll
    for (int k = M - 3; k >= 0; k--) r[k] = fma(a1, r[k + 1], fma(a2, r[k + 2], r[k]));
    double b0 = r[0], b1 = M > 1 ? r[1] : 0.0;
    {
        int j = 0;
        for (int d = 1; d < P; d <<= 1, j++) {
            const double t0 = __shfl(b0, L.src(d)), t1 = __shfl(b1, L.src(d));
            if (c + d < P) {
                double y0, y1;
                mv(T.sc[j], t0, t1, y0, y1);
                b0 += y0;
                b1 += y1;
            }
        }
    }
    w0 = __shfl(b0, L.seg);
    w1 = __shfl(b1, L.seg);
    mv(T.wrap, w0, w1, tot0, tot1);
    p0 = __shfl(b0, L.src(1));
    p1 = __shfl(b1, L.src(1));
    if (c == P - 1) { p0 = 0.0; p1 = 0.0; }
    mv(T.lp[P - 1 - c], tot0, tot1, q0, q1);
    const double bin0 = p0 + q0, bin1 = p1 + q1;
#pragma unroll
    for (int k = 0; k < M; k++) r[k] = fma(T.cf[M - 1 - k][0], bin0, fma(T.cf[M - 1 - k][1], bin1, r[k]));
}

// ---- operators on a window -----------------------------------------------

// First derivative: K = cadd*sx(1)*(N-1)*(1+rho^2).
template <int M>
DEV void op_d1(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((C.c1 * (w[i + 3] - w[i - 3]) + C.b1 * (w[i + 2] - w[i - 2])) + C.a1 * (w[i + 1] - w[i - 1]));
    }
    solve1<M>(out, T, L);
}

// Second derivative: K = (N-1)^2*sx(2)*(1+rho^2).
template <int M>
DEV void op_d2(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * (((C.c2 * (w[i + 3] + w[i - 3]) + C.b2 * (w[i + 2] + w[i - 2])) + C.a2 * (w[i + 1] + w[i - 1])) - C.d2 * w[i]);
    }
    solve1<M>(out, T, L);
}

// Eighth-order compact filter: K = 1/kappa.
template <int M>
DEV void op_filt(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
                 const Lane<M> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((((C.f4 * (w[i + 4] + w[i - 4]) + C.f3 * (w[i + 3] + w[i - 3])) + C.f2 * (w[i + 2] + w[i - 2]))
                       + C.f1 * (w[i + 1] + w[i - 1])) + C.f0 * w[i]);
    }
    solve2<M>(out, T, L);
}

// Truncated Gaussian (explicit).
template <int M>
DEV void op_gauss(const double (&w)[M + 2 * HALO], const Coef &C, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = (((C.g4 * (w[i + 4] + w[i - 4]) + C.g3 * (w[i + 3] + w[i - 3])) + C.g2 * (w[i + 2] + w[i - 2]))
                  + C.g1 * (w[i + 1] + w[i - 1])) + C.g0 * w[i];
    }
}

// ---- pointwise thermodynamics (src/rhsv_periodic_sf.f:92-104,237-250) ----

struct Prim {
    double vs, u, v, w, p, T;
};

DEV Prim prim(const Phys &ph, double q1, double q2, double q3, double q4, double q5)
{
    Prim r;
    r.vs = 1.0 / q4;
    const double ke = q1 * q1 + q2 * q2 + q3 * q3;
    r.p = ph.gm1 * (q5 - 0.5 * r.vs * ke);
    r.T = ph.gfac * r.vs * r.p;
    r.u = q1 * r.vs;
    r.v = q2 * r.vs;
    r.w = q3 * r.vs;
    return r;
}

DEV double visc(const Phys &ph, double T)
{
    return pow(ph.gm1 * T, ph.vis_exp) * ph.re_inv;
}

DEV double dvisc(const Phys &ph, double T)
{
    return pow(ph.gm1 * T, ph.vis_exp_m1) * ph.re_inv * ph.vis_exp * ph.gm1;
}

}  // namespace st
