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
    long ntiles;           // over all slots of the launch
    long tpi;              // tiles per instance (slot)
};

// Tile for this workgroup.  Workgroups are dispatched round-robin over the
// 8 XCDs; consecutive tiles (which share 128-byte lines for the strided
// directions) are handed to workgroups of the same XCD.
DEV long tile_of(const Geo &g, long wg)
{
    if ((g.ntiles & 7) == 0) return (wg & 7) * (g.ntiles >> 3) + (wg >> 3);
    return wg;
}

// LDS layout of one slot: line b holds its n = M*P points in a block of
// n doubles; point s = c*M + k sits at row k, column c.  Rows are rotated by
// b and columns XOR-swizzled by a key f(k, b).  The LDS serves 16 lanes of a
// 64-bit access per cycle over 16 double-wide banks, so the key is chosen
// such that every group of 16 lanes hits 16 distinct banks in all three
// access patterns: the line-order reads and writes of the compute phase
// (fixed k, consecutive c), the memory-order staging of a y tile
// (consecutive s: 16/M chunks by M rows) and that of an x or z tile (B lines
// side by side, 16/B points of each).  For P >= 16 (row offsets are then
// multiples of 16) f = (16/M) k + g(b): the point-in-chunk bits sit above
// the chunk bits of the y staging, and lds_g places the line's bits where
// the x/z staging leaves the 16 banks free.  Narrower lines (P = 8) keep
// f = k (P/M) + b (4b for M = 4).  (M = 3 or 6: rows rotate by b mod M; any
// rotation and swizzle is a bijection of the block, only the bank spread
// depends on them.)
// The additive layout (AL, M a power of two) drops the rotation and the
// row-dependent part of the key: point s = c*M + k of line b sits at
// b*n + k*P + (c ^ g(b)).  A lane's points then lie at one base plus
// constant offsets, so one address register serves all of its rows (the
// swizzled layout needs one per row, which the passes with the most live
// values spill).  The compute phase and the staging of x and z tiles of 16
// or more lines stay free of bank conflicts; the staging of y tiles and of
// narrower x and z tiles takes up to M-way conflicts.  Each pass names the
// layout that measured faster for it (LDS_ADD).
template <int M>
HD int lds_g(int b, int B)
{
    int lB = 0;
    while ((1 << lB) < B) lB++;
    constexpr int lM = M >= 8 ? 3 : M >= 4 ? 2 : M >= 2 ? 1 : 0;
    if (lB + lM <= 4) return b << (4 - lB - lM);
    const int lo = 4 - lM, kb = lB < 4 ? 4 - lB : 0;
    return (b & ((1 << lo) - 1)) | ((b >> lo) << (lo + kb));
}

template <int M, bool AL = false>
HD int lds_idx(int b, int k, int c, int P, int g)
{
    if constexpr ((M & (M - 1)) == 0) {
        if constexpr (AL) return b * (M * P) + k * P + (c ^ (g & (P - 1)));
        const int f = P >= 16 ? (16 / M) * k + g : k * (P / M) + (M == 4 ? 4 * b : b);
        return b * (M * P) + (((k + b) * P + (c ^ (f & (P - 1)))) & (M * P - 1));
    } else {
        int r = k + b % M;
        r = r >= M ? r - M : r;
        return b * (M * P) + r * P + (c ^ ((k + b) & (P - 1)));
    }
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
    int t;                // the thread

    // tile: the tile within its instance; off: the instance's first point;
    // tid: the thread index (threadIdx.x, or a copy the compiler cannot see
    // through, so that a second walk is computed again instead of kept)
    DEV Walk(const Geo &g, long tile, long off, int tid)
    {
        t = tid;
        const long line0 = tile * B;
        if (g.dir == 1) {
            base = off + line0 * n + t;
            step = NT;
            b = 0;
            s0 = 0;
            return;
        }
        b = t & (B - 1);
        s0 = t / B;
        if (g.dir == 0) {
            const long sx = (long)g.My * g.Mz;
            base = off + line0 + b + s0 * sx;
            step = P * sx;
        } else {
            const unsigned line = (unsigned)(line0 + b);
            const unsigned i = line / (unsigned)g.My, j = line - i * (unsigned)g.My;
            base = off + j + (long)g.My * g.Mz * i + (long)s0 * g.My;
            step = (long)P * g.My;
        }
    }
    // global index and (line, point) of the r-th element
    DEV long at(const Geo &g, int r, int &bb, int &ss) const
    {
        if (g.dir == 1) {
            const int e = t + NT * r;
            bb = e / n;
            ss = e - bb * n;
        } else {
            bb = b;
            ss = s0 + r * P;
        }
        return base + r * step;
    }
};

// Per-lane view during the compute phase; P lanes per line.  NP: the line
// is non-periodic (closure rows at both ends, no wrap).
template <int M, int P_, bool NP_ = false, bool AL_ = false>
struct Lane {
    static constexpr int P = P_;
    static constexpr bool NP = NP_;
    static constexpr bool AL = AL_;   // additive LDS layout
    int c, seg, b;         // chunk index, first lane of the line, line in tile
    int g;                 // the line's column key (lds_g)
    double *lds;
    int ss;                // slot size (doubles)

    DEV double ld(int slot, int k) const { return lds[slot * ss + lds_idx<M, AL>(b, k, c, P, g)]; }
    DEV void st(int slot, int k, double v) const { lds[slot * ss + lds_idx<M, AL>(b, k, c, P, g)] = v; }
    DEV int src(int dc) const { return seg + ((c + dc) & (P - 1)); }
};

// ss = B*M*P = NT*M doubles: a compile-time constant for the kernel, so the
// slot offsets fold into the LDS instructions' immediate offsets (and, with
// the additive layout, the row offsets too).
template <int M, int P, bool NP = false, bool AL = false>
DEV Lane<M, P, NP, AL> make_lane(double *lds, int ss)
{
    Lane<M, P, NP, AL> L;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    L.c = lane & (P - 1);
    L.seg = lane - L.c;
    L.b = wave * (64 / P) + lane / P;
    L.g = lds_g<M>(L.b, ss / (M * P));
    L.lds = lds;
    L.ss = ss;
    return L;
}

// Window of M own values plus HALO on each side, fetched from neighbours.
// On a non-periodic line the points beyond either end are the mirror
// images f(-m) = f(m), f(n-1+m) = f(n-1-m) (the symmetry of
// src/filters_nonperiodic1d.f:830-943); the compact operators overwrite
// their closure rows and never read them.
template <int M, int P, bool NP, bool AL>
DEV void window(const double (&v)[M], const Lane<M, P, NP, AL> &L, double (&w)[M + 2 * HALO])
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
    if constexpr (NP) {
        constexpr int n = M * P;
        // chunks whose window reaches past the first point: row -j sits
        // at w[HALO - c*M - j], its image row j at w[HALO - c*M + j]
        // (selects, not branches: every index stays a constant)
#pragma unroll
        for (int cc = 0; cc * M < HALO; cc++) {
            const bool here = L.c == cc;
#pragma unroll
            for (int j = 1; j <= HALO; j++) {
                const int at = HALO - cc * M - j;
                if (at >= 0) w[at] = here ? w[HALO - cc * M + j] : w[at];
            }
        }
        // and past the last: row n-1+j at w[HALO + n-1+j - c*M]
#pragma unroll
        for (int cc = P - 1; cc >= 0 && cc * M + M - 1 + HALO > n - 1; cc--) {
            const bool here = L.c == cc;
#pragma unroll
            for (int j = 1; j <= HALO; j++) {
                const int at = HALO + n - 1 + j - cc * M;
                if (at < M + 2 * HALO) w[at] = here ? w[HALO + n - 1 - j - cc * M] : w[at];
            }
        }
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
template <int K, int M, int P, bool NP, bool AL>
DEV void solve(double (&r)[M], const OpTab &T, const Lane<M, P, NP, AL> &L)
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
    if constexpr (!NP) {
        madd<K>(T.wrap, tot, w);
        madd<K>(T.lp[c], w, cin);
    }
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
    if constexpr (!NP) {
        madd<K>(T.wrap, tot, w);
        madd<K>(T.lp[P - 1 - c], w, cin);
    }
#pragma unroll
    for (int k = 0; k < M; k++) {
        r[k] = fma(T.cf[M - 1 - k][0], cin[0], r[k]);
        if (K == 2) r[k] = fma(T.cf[M - 1 - k][1], cin[K - 1], r[k]);
    }
}

// ---- non-periodic lines ---------------------------------------------------

// Value at row ROW of this lane's line after a solve (every lane gets it).
template <int M, int P, bool NP, bool AL>
DEV double line_value(const double (&x)[M], const Lane<M, P, NP, AL> &L, int row)
{
    const int src = row / M;
    if constexpr (P == 64) {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x[row % M]), src),
                                __builtin_amdgcn_readlane(__double2loint(x[row % M]), src));
    } else {
        return __shfl(x[row % M], L.seg + src);
    }
}

// Boundary correction x -= Z q, q = G yb (the np_* tables of OpTab; Z
// itself is applied through the recurrences, below).
template <int MB, int M, int P, bool NP, bool AL>
DEV void np_fix(double (&x)[M], const OpTab &T, const Lane<M, P, NP, AL> &L)
{
    constexpr int n = M * P;
    double yb[10];
#pragma unroll
    for (int t = 0; t < 5; t++) {
        yb[t] = line_value<M, P>(x, L, t);
        yb[5 + t] = line_value<M, P>(x, L, n - 5 + t);
    }
    double q[MB];
#pragma unroll
    for (int j = 0; j < MB; j++) {
        double s = T.npG[j][0] * yb[0];
#pragma unroll
        for (int t = 1; t < 10; t++) s = fma(T.npG[j][t], yb[t], s);
        q[j] = s;
    }
    // Z q = U^-1 L^-1 (scale sum_j q_j e_(row j)), the boundary rows being
    // the first and last MB/2 of the line: one more solve of a right-hand
    // side that is zero away from them, instead of the MB columns of Z per
    // point (which, read once per kernel, would hold MB*M registers per
    // operator, and read at each use, wait on memory in every solve)
    constexpr int K = MB == 6 ? 2 : 1, H = MB / 2;
    double z[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = L.c * M + k;
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < H; j++) {
            v = i == j ? T.scale * q[j] : v;
            v = i == n - H + j ? T.scale * q[H + j] : v;
        }
        z[k] = v;
    }
    solve<K, M>(z, T, L);
#pragma unroll
    for (int k = 0; k < M; k++) x[k] -= z[k];
}

// Closure rows of the right-hand side.  OP 0: first derivative
// (src/derivCDopt6_nonperiodic1d.f:463-471, 494-502, with the closure
// constants of src/header:173-175), 1: second derivative (:615-629, 653-668),
// 2: compact filter (src/filters_nonperiodic1d.f:324-349, 384-408, with
// src/LES_header:118-148).  f(j) reads the j-th point from the near end;
// the first-derivative rows at the far end are the negated mirror images.
// The one-sided rows cancel strongly (the second-derivative row 1 sums
// |coefficients| 56 times its result), so they keep the reference's
// rounding: every product rounded, sums left to right, no contraction.
template <int OP, class F>
DEV double closure_row(int r, const F &f)
{
#pragma clang fp contract(off)
    if constexpr (OP == 0) {
        constexpr double cl1 = 5.0 / 9.0, cl2 = 1.0 / 18.0, cl3 = 1.0 / 36.0, cl4 = 7.0 / 9.0;
        if (r == 0) return (-2.5 * f(0) + 2.0 * f(1)) + 5.0e-1 * f(2);
        if (r == 1) return ((-cl1 * f(0) - 5.0e-1 * f(1)) + f(2)) + cl2 * f(3);
        return cl3 * (f(4) - f(0)) + cl4 * (f(3) - f(1));
    } else if constexpr (OP == 1) {
        constexpr double cl5 = 3.0 / 44.0, cl6 = 12.0 / 11.0;
        if (r == 0) return ((1.3e1 * f(0) - 2.7e1 * f(1)) + 1.5e1 * f(2)) - f(3);
        if (r == 1) return (((1.2375 * f(0) - 3.0 * f(1)) + 2.325 * f(2)) - 6.0e-1 * f(3)) + 3.75e-2 * f(4);
        return cl5 * ((f(0) - 2.0 * f(2)) + f(4)) + cl6 * ((f(1) - 2.0 * f(2)) + f(3));
    } else {
        constexpr double ab = 0.45, al = 6.6624e-01, be = 1.6688e-01;
        if (r == 0)
            return (((15.0 / 16.0 * f(0) + 1.0 / 4.0 * f(1)) + -3.0 / 8.0 * f(2)) + 1.0 / 4.0 * f(3)) + -1.0 / 16.0 * f(4);
        if (r == 1) {
            constexpr double c1 = 1.0 / 16.0 + (7.0 / 8.0) * ab, c2 = 3.0 / 4.0 + (1.0 / 2.0) * ab;
            constexpr double c3 = 3.0 / 8.0 + (1.0 / 4.0) * ab, c4 = -1.0 / 4.0 + (1.0 / 2.0) * ab;
            constexpr double c5 = 1.0 / 16.0 - (1.0 / 8.0) * ab;
            return (((c1 * f(0) + c2 * f(1)) + c3 * f(2)) + c4 * f(3)) + c5 * f(4);
        }
        if (r == 2) {
            constexpr double c0 = 5.0 / 8.0 + (3.0 / 4.0) * ab, c1 = (1.0 / 2.0 + ab) / 2.0;
            constexpr double c2 = (-1.0 / 8.0 + ab / 4.0) / 2.0;
            return (((c2 * f(4) + c1 * f(3)) + c0 * f(2)) + c1 * f(1)) + c2 * f(0);
        }
        constexpr double c0 = 11.0 / 16.0 + (5.0 / 8.0) * al - (5.0 / 8.0) * be;
        constexpr double c1 = (15.0 / 32.0 + (17.0 / 16.0) * al + (15.0 / 16.0) * be) / 2.0;
        constexpr double c2 = (-3.0 / 16.0 + (3.0 / 8.0) * al + (13.0 / 8.0) * be) / 2.0;
        constexpr double c3 = (1.0 / 32.0 - al / 16.0 + be / 16.0) / 2.0;
        return (((((c3 * f(6) + c2 * f(5)) + c1 * f(4)) + c0 * f(3)) + c1 * f(2)) + c2 * f(1)) + c3 * f(0);
    }
}

// Overwrite the closure rows (the first NB and last NB of the line) of the
// right-hand side out = K * stencil.  The chunks that hold them are
// enumerated at compile time so every window index is a constant.
template <int OP, int M, int P, bool NP, bool AL>
DEV void closure_rows(const double (&w)[M + 2 * HALO], double K, const Lane<M, P, NP, AL> &L, double (&out)[M])
{
    constexpr int n = M * P, NB = OP == 2 ? 4 : 3;
    // (selects, not branches: every index stays a constant)
#pragma unroll
    for (int cc = 0; cc * M < NB; cc++) {
        const bool here = L.c == cc;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const int i = cc * M + k;
            if (i < NB) {
                const int o = HALO - cc * M;
                const double v = closure_row<OP>(i, [&](int j) { return w[o + j]; });
                out[k] = here ? K * v : out[k];
            }
        }
    }
#pragma unroll
    for (int cc = P - 1; cc >= 0 && cc * M + M > n - NB; cc--) {
        const bool here = L.c == cc;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const int i = cc * M + k;
            if (i >= n - NB) {
                const int o = HALO + n - 1 - cc * M;
                const double v = closure_row<OP>(n - 1 - i, [&](int j) { return w[o - j]; });
                out[k] = here ? K * (OP == 0 ? -v : v) : out[k];
            }
        }
    }
}

// ---- operators on a window -----------------------------------------------

// First derivative: K = cadd*sx(1)*(N-1)*(1+rho^2).
template <int M, int P, bool NP, bool AL>
DEV void op_d1(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M, P, NP, AL> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((C.c1 * (w[i + 3] - w[i - 3]) + C.b1 * (w[i + 2] - w[i - 2])) + C.a1 * (w[i + 1] - w[i - 1]));
    }
    if constexpr (NP) closure_rows<0>(w, K, L, out);
    solve<1, M>(out, T, L);
    if constexpr (NP) np_fix<4>(out, T, L);
}

// Second derivative: K = (N-1)^2*sx(2)*(1+rho^2).
template <int M, int P, bool NP, bool AL>
DEV void op_d2(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
               const Lane<M, P, NP, AL> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * (((C.c2 * (w[i + 3] + w[i - 3]) + C.b2 * (w[i + 2] + w[i - 2])) + C.a2 * (w[i + 1] + w[i - 1])) - C.d2 * w[i]);
    }
    if constexpr (NP) closure_rows<1>(w, K, L, out);
    solve<1, M>(out, T, L);
    if constexpr (NP) np_fix<4>(out, T, L);
}

// Eighth-order compact filter: K = 1/kappa.
template <int M, int P, bool NP, bool AL>
DEV void op_filt(const double (&w)[M + 2 * HALO], const Coef &C, double K, const OpTab &T,
                 const Lane<M, P, NP, AL> &L, double (&out)[M])
{
#pragma unroll
    for (int k = 0; k < M; k++) {
        const int i = k + HALO;
        out[k] = K * ((((C.f4 * (w[i + 4] + w[i - 4]) + C.f3 * (w[i + 3] + w[i - 3])) + C.f2 * (w[i + 2] + w[i - 2]))
                       + C.f1 * (w[i + 1] + w[i - 1])) + C.f0 * w[i]);
    }
    if constexpr (NP) closure_rows<2>(w, K, L, out);
    solve<2, M>(out, T, L);
    if constexpr (NP) np_fix<6>(out, T, L);
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

// Rounded as the reference rounds it (no contraction): the temperature
// feeds second derivatives whose one-sided boundary rows cancel strongly,
// so an ulp of difference here would show up there.
DEV Prim prim(const Phys &ph, double q1, double q2, double q3, double q4, double q5)
{
#pragma clang fp contract(off)
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

// Fix_Pressure at a point (src/lddrk4_periodic_sf.f:436-473): the energy
// of the state (q1..q5) with P = (gamma-1)(q5 - |q1..3|^2 / 2 q4) clamped to
// 0.71463 where negative.  Without contraction, so that every pass that
// applies it rounds alike.
DEV double fixed_energy(const Phys &ph, double q1, double q2, double q3, double q4, double q5)
{
#pragma clang fp contract(off)
    const double vs = 1.0 / q4;
    const double ke = q1 * q1 + q2 * q2 + q3 * q3;
    double p = ph.gm1 * (q5 - 0.5 * vs * ke);
    if (p < 0.0) p = 0.71463;
    return ph.gm1_inv * p + 0.5 * vs * ke;
}

// Compute_CFL's per-point spectral radii (src/lddrk4_periodic_sf.f:29-49):
// sum over directions of (c + |u_d|) k / Delta_d with k = 2.478 (w1, the
// stability bound) and 2.199 (w2, the accuracy bound), c = sqrt(gamma (gamma-1) e).
DEV void cfl_point(const Phys &ph, double q1, double q2, double q3, double q4, double q5, double &w1, double &w2)
{
    const double fac = ph.gamma * ph.gm1;
    const double t1 = 1.0 / q4;
    const double t2 = sqrt(fac * t1 * (q5 - 0.5 * t1 * (q1 * q1 + q2 * q2 + q3 * q3)));
    const double ux = fabs(q1 * t1), uy = fabs(q2 * t1), uz = fabs(q3 * t1);
    w1 = (t2 + ux) * 2.478 / ph.D[0] + (t2 + uy) * 2.478 / ph.D[1] + (t2 + uz) * 2.478 / ph.D[2];
    w2 = (t2 + ux) * 2.199 / ph.D[0] + (t2 + uy) * 2.199 / ph.D[1] + (t2 + uz) * 2.199 / ph.D[2];
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
