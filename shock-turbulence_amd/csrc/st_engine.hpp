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
#include "st_math.h"

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

// Memory-order walk of one tile by one thread.  The tile holds B = NT/P
// lines of n = M*P points; the thread visits elements e = tid + NT*r,
// r = 0..M-1.  x and z tiles are B adjacent lines (consecutive y, the
// unit-stride direction), so a thread stays on one line b = tid mod B and
// steps s = tid/B + r*P along it; a y tile is B*n contiguous points.
template <int M, int P, int NT>
struct Walk {
    static constexpr int n = M * P, B = NT / P;
    long base;            // global index of the thread's first element
    long step;            // global index step per r
    int b, s0;            // x and z: the thread's line and first point

    DEV Walk(const Geo &g, long tile)
    {
        const int t = threadIdx.x;
        const long line0 = tile * B;
        if (g.dir == 1) {
            base = line0 * n + t;
            step = NT;
            b = 0;
            s0 = 0;
            return;
        }
        b = t & (B - 1);
        s0 = t / B;
        if (g.dir == 0) {
            const long sx = (long)g.My * g.Mz;
            base = line0 + b + s0 * sx;
            step = P * sx;
        } else {
            const unsigned line = (unsigned)(line0 + b);
            const unsigned i = line / (unsigned)g.My, j = line - i * (unsigned)g.My;
            base = j + (long)g.My * g.Mz * i + (long)s0 * g.My;
            step = (long)P * g.My;
        }
    }
    // global index and (line, point) of the r-th element
    DEV long at(const Geo &g, int r, int &bb, int &ss) const
    {
        if (g.dir == 1) {
            const int e = threadIdx.x + NT * r;
            bb = e / n;
            ss = e & (n - 1);
        } else {
            bb = b;
            ss = s0 + r * P;
        }
        return base + r * step;
    }
};

// Per-lane view during the compute phase; P lanes per line.
template <int M, int P_>
struct Lane {
    static constexpr int P = P_;
    int c, seg, b;         // chunk index, first lane of the line, line in tile
    double *lds;
    int ss;                // slot size (doubles)

    DEV double ld(int slot, int k) const { return lds[slot * ss + lds_idx<M>(b, k, c, P)]; }
    DEV void st(int slot, int k, double v) const { lds[slot * ss + lds_idx<M>(b, k, c, P)] = v; }
    DEV int src(int dc) const { return seg + ((c + dc) & (P - 1)); }
};

// ss = B*M*P = NT*M doubles: a compile-time constant for the kernel, so the
// slot offsets fold into the LDS instructions' immediate offsets.
template <int M, int P>
DEV Lane<M, P> make_lane(double *lds, int ss)
{
    Lane<M, P> L;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    L.c = lane & (P - 1);
    L.seg = lane - L.c;
    L.b = wave * (64 / P) + lane / P;
    L.lds = lds;
    L.ss = ss;
    return L;
}

// Window of M own values plus HALO on each side, fetched from neighbours.
template <int M, int P>
DEV void window(const double (&v)[M], const Lane<M, P> &L, double (&w)[M + 2 * HALO])
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
//
// The lane-to-lane traffic of a solve runs on DPP (data-parallel
// primitives, a VALU operand modifier) and v_readlane instead of LDS
// permutes: a scan step costs a few VALU cycles instead of an LDS round trip.

constexpr int DPP_ROW_SHL = 0x100;      // + d: lane i <- lane i+d within its 16-lane row
constexpr int DPP_ROW_SHR = 0x110;      // + d: lane i <- lane i-d within its row
constexpr int DPP_WAVE_SHL1 = 0x130;    // lane i <- lane i+1 across the wave
constexpr int DPP_WAVE_SHR1 = 0x138;    // lane i <- lane i-1 across the wave
constexpr int DPP_ROW_NEWBCAST = 0x150; // + k: every lane of a row <- lane k of the row

// Lanes whose source lies outside the row (or the wave) read 0.
template <int CTRL>
DEV double dpp(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

template <int LANE>
DEV double rdlane(double x)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), LANE),
                            __builtin_amdgcn_readlane(__double2loint(x), LANE));
}

// y += m t for the carried state: a scalar (order 1, m[0]) or a 2-vector
// with the row-major 2x2 matrix m (order 2).
template <int K>
DEV void madd(const double *m, const double (&t)[K], double (&y)[K])
{
    if (K == 1) {
        y[0] = fma(m[0], t[0], y[0]);
    } else {
        const double t1 = t[K - 1];
        y[0] = fma(m[0], t[0], fma(m[1], t1, y[0]));
        y[K - 1] = fma(m[2], t[0], fma(m[3], t1, y[K - 1]));
    }
}

template <int K>
DEV void madd_to(const double *m, const double (&t)[K], const double (&x)[K], double (&y)[K])
{
#pragma unroll
    for (int i = 0; i < K; i++) y[i] = x[i];
    madd<K>(m, t, y);
}

// One in-row Kogge-Stone step at distance D: forward e_c += A^(M D) e_(c-D),
// backward e_c += A^(M D) e_(c+D).  Rows are segment boundaries for P >= 16;
// for P = 8 two lines share a row and the cross-line source is masked.
template <int K, bool FWD, int D, int J, int P>
DEV void row_step(double (&e)[K], const OpTab &T, int c)
{
    double t[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        t[i] = dpp<FWD ? DPP_ROW_SHR + D : DPP_ROW_SHL + D>(e[i]);
        if constexpr (P < 16) {
            if (FWD ? c < D : c + D >= P) t[i] = 0.0;
        }
    }
    madd<K>(T.sc[J], t, e);
}

template <int K>
DEV void sel(bool p, const double (&a)[K], const double (&b)[K], double (&y)[K])
{
#pragma unroll
    for (int i = 0; i < K; i++) y[i] = p ? a[i] : b[i];
}

// Segmented scan of the chunk-end states e over the P lanes of each line.
// On return e is the inclusive prefix (forward) or suffix (backward) state,
// and tot the line's total (the state at its last, resp. first, lane).
template <int K, bool FWD, int P>
DEV void wave_scan(double (&e)[K], double (&tot)[K], const OpTab &T, int c)
{
    static_assert(P == 8 || P == 16 || P == 32 || P == 64, "lanes per line");
    row_step<K, FWD, 1, 0, P>(e, T, c);
    row_step<K, FWD, 2, 1, P>(e, T, c);
    row_step<K, FWD, 4, 2, P>(e, T, c);
    if constexpr (P == 8) {
        // two lines per row: totals at row lanes 7 / 15 (forward), 0 / 8 (backward)
        const bool hi = (threadIdx.x & 8) != 0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            const double a = FWD ? dpp<DPP_ROW_NEWBCAST + 15>(e[i]) : dpp<DPP_ROW_NEWBCAST + 8>(e[i]);
            const double b = FWD ? dpp<DPP_ROW_NEWBCAST + 7>(e[i]) : dpp<DPP_ROW_NEWBCAST + 0>(e[i]);
            tot[i] = hi ? a : b;
        }
    } else {
        row_step<K, FWD, 8, 3, P>(e, T, c);
        if constexpr (P == 16) {
#pragma unroll
            for (int i = 0; i < K; i++)
                tot[i] = FWD ? dpp<DPP_ROW_NEWBCAST + 15>(e[i]) : dpp<DPP_ROW_NEWBCAST + 0>(e[i]);
        } else {
            // P = 32 or 64: carry whole rows with uniform values read from
            // one lane per row (its row-local end state), A16 = A^(16 M).
            const double *A16 = T.sc[4];
            const int row = (threadIdx.x >> 4) & 3;
            double s0[K], s1[K], s2[K], s3[K], carry[K];
#pragma unroll
            for (int i = 0; i < K; i++) {
                s0[i] = FWD ? rdlane<15>(e[i]) : rdlane<0>(e[i]);
                s1[i] = FWD ? rdlane<31>(e[i]) : rdlane<16>(e[i]);
                s2[i] = FWD ? rdlane<47>(e[i]) : rdlane<32>(e[i]);
                s3[i] = FWD ? rdlane<63>(e[i]) : rdlane<48>(e[i]);
            }
            const double zero[K] = {};
            if constexpr (P == 32) {
                double t0[K], t1[K], c0[K];
                if (FWD) {
                    madd_to<K>(A16, s0, s1, t0);           // total of lanes 0..31
                    madd_to<K>(A16, s2, s3, t1);           // total of lanes 32..63
                    sel<K>(row == 1, s0, zero, c0);
                    sel<K>(row == 3, s2, c0, carry);
                } else {
                    madd_to<K>(A16, s1, s0, t0);
                    madd_to<K>(A16, s3, s2, t1);
                    sel<K>(row == 0, s1, zero, c0);
                    sel<K>(row == 2, s3, c0, carry);
                }
                sel<K>(row >= 2, t1, t0, tot);
            } else {
                double f1[K], f2[K], c0[K], c1[K];
                if (FWD) {
                    madd_to<K>(A16, s0, s1, f1);           // prefix through lane 31
                    madd_to<K>(A16, f1, s2, f2);           // through lane 47
                    madd_to<K>(A16, f2, s3, tot);          // through lane 63
                    sel<K>(row == 1, s0, zero, c0);
                    sel<K>(row == 2, f1, c0, c1);
                    sel<K>(row == 3, f2, c1, carry);
                } else {
                    madd_to<K>(A16, s3, s2, f2);           // suffix from lane 32
                    madd_to<K>(A16, f2, s1, f1);           // from lane 16
                    madd_to<K>(A16, f1, s0, tot);          // from lane 0
                    sel<K>(row == 0, f1, zero, c0);
                    sel<K>(row == 1, f2, c0, c1);
                    sel<K>(row == 2, s3, c1, carry);
                }
            }
            // the carry enters lane c over (c mod 16) + 1 chunks (forward)
            // or 16 - (c mod 16) chunks (backward)
            madd<K>(T.lp[FWD ? (c & 15) + 1 : 16 - (c & 15)], carry, e);
        }
    }
}

// Cyclic solve of one line per P lanes; r holds the lane's M right-hand
// sides (the scale 1+rho^2 or 1/kappa already folded in) and returns x.
// Order 1: x = (1 - rho S^-1)^-1 (1 - rho S)^-1 r,
// order 2: x = (1 - a1 S^-1 - a2 S^-2)^-1 (1 - a1 S - a2 S^2)^-1 r.
template <int K, int M, int P>
DEV void solve(double (&r)[M], const OpTab &T, const Lane<M, P> &L)
{
    const double a1 = T.a1, a2 = T.a2;
    const int c = L.c;
    // forward recurrence inside the chunk (zero incoming state)
#pragma unroll
    for (int k = 1; k < M; k++) {
        r[k] = fma(a1, r[k - 1], r[k]);
        if (K == 2 && k >= 2) r[k] = fma(a2, r[k - 2], r[k]);
    }
    double e[K], tot[K], w[K], cin[K];
    e[0] = r[M - 1];
    if (K == 2) e[K - 1] = M > 1 ? r[M > 1 ? M - 2 : 0] : 0.0;
    wave_scan<K, true, P>(e, tot, T, c);
    // incoming state of chunk c: the exclusive prefix plus the cyclic total
    // carried around the line, A^(M c) (I - A^n)^-1 tot
#pragma unroll
    for (int i = 0; i < K; i++) {
        cin[i] = dpp<DPP_WAVE_SHR1>(e[i]);
        if (c == 0) cin[i] = 0.0;
        w[i] = 0.0;
    }
    madd<K>(T.wrap, tot, w);
    madd<K>(T.lp[c], w, cin);
#pragma unroll
    for (int k = 0; k < M; k++) {
        r[k] = fma(T.cf[k][0], cin[0], r[k]);
        if (K == 2) r[k] = fma(T.cf[k][1], cin[K - 1], r[k]);
    }
    // backward recurrence inside the chunk
#pragma unroll
    for (int k = M - 2; k >= 0; k--) {
        r[k] = fma(a1, r[k + 1], r[k]);
        if (K == 2 && k <= M - 3) r[k] = fma(a2, r[k + 2], r[k]);
    }
    e[0] = r[0];
    if (K == 2) e[K - 1] = M > 1 ? r[M > 1 ? 1 : 0] : 0.0;
    wave_scan<K, false, P>(e, tot, T, c);
#pragma unroll
    for (int i = 0; i < K; i++) {
        cin[i] = dpp<DPP_WAVE_SHL1>(e[i]);
        if (c == P - 1) cin[i] = 0.0;
        w[i] = 0.0;
    }
    madd<K>(T.wrap, tot, w);
    madd<K>(T.lp[P - 1 - c], w, cin);
#pragma unroll
    for (int k = 0; k < M; k++) {
        r[k] = fma(T.cf[M - 1 - k][0], cin[0], r[k]);
        if (K == 2) r[k] = fma(T.cf[M - 1 - k][1], cin[K - 1], r[k]);
    }
}

// ---- operators on a window -----------------------------------------------

// First derivative: K = cadd*sx(1)*(N-1)*(1+rho^2).
template <int M, int P>
DEV void op_d1(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M, P> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((C.c1 * (w[i + 3] - w[i - 3]) + C.b1 * (w[i + 2] - w[i - 2])) + C.a1 * (w[i + 1] - w[i - 1]));
    }
    solve<1, M>(out, T, L);
}

// Second derivative: K = (N-1)^2*sx(2)*(1+rho^2).
template <int M, int P>
DEV void op_d2(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M, P> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * (((C.c2 * (w[i + 3] + w[i - 3]) + C.b2 * (w[i + 2] + w[i - 2])) + C.a2 * (w[i + 1] + w[i - 1])) - C.d2 * w[i]);
    }
    solve<1, M>(out, T, L);
}

// Eighth-order compact filter: K = 1/kappa.
template <int M, int P>
DEV void op_filt(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
                 const Lane<M, P> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((((C.f4 * (w[i + 4] + w[i - 4]) + C.f3 * (w[i + 3] + w[i - 3])) + C.f2 * (w[i + 2] + w[i - 2]))
                       + C.f1 * (w[i + 1] + w[i - 1])) + C.f0 * w[i]);
    }
    solve<2, M>(out, T, L);
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

// mu = (gamma-1 T)^vis_exp / Re and d mu/dT = vis_exp mu / T
// (src/rhsv_periodic_sf.f:245-247, the derivative written without a second
// power; the two forms differ by rounding).
DEV double visc(const Phys &ph, double T)
{
    return series_pow(ph.gm1 * T, ph.vis_exp) * ph.re_inv;
}

DEV double dvisc(const Phys &ph, double T, double mu)
{
    return ph.vis_exp * mu / T;
}

}  // namespace st
