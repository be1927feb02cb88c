// HIP kernels for the per-RK-stage pipeline (gfx950, FP64).
//
// The reference evaluates one stage as ~110 separate compact operators,
// each copying bands, transposing and re-eliminating
// (src/derivCDopt6_periodic.f, src/pentadiagonal.f90).  Here a stage is a
// short chain of directional passes; each pass loads its inputs once,
// applies every operator of that direction the dataflow allows, folds the
// pointwise algebra of src/rhsv_periodic_sf.f and src/hyper_rhs.f into
// its load and store phases, and writes its results once.
#include <hip/hip_runtime.h>

#include "st_engine.hpp"
#include "st_kernels.h"

namespace st {

// ---------------------------------------------------------------------
// generic pass kernel
// ---------------------------------------------------------------------
template <class PASS>
constexpr int pass_slots() { return PASS::REUSE ? PASS::NSLOT : PASS::NSLOT + PASS::NEARLY; }

// Slots 0..NSLOT-1 hold the staged inputs; compute() may write final
// results early (freeing their registers) into slots NSLOT..NSLOT+NEARLY-1,
// or, with REUSE, over its own points of the input slots 0..NEARLY-1 once
// it no longer reads them (each lane reads and writes only its own points
// of the tile during compute).  The NOUT register results are staged into
// slots 0..NOUT-1 after compute.  Without REUSE, store() receives per point
// the NOUT late results followed by the NEARLY early ones; with REUSE the
// early ones go first through store_early() and store() gets the late ones.
// A pass may also name NPF values per point that its store phase reads
// (read-modify-write operands): prefetch() loads them during the load phase,
// which visits the same points per thread as the store phase, and they
// reach store_early() (REUSE) or store() appended after the results.
template <int M, int P, int NT, class PASS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PASS::MIN_WAVES)))
k_pass(Geo g, PassArgs a)
{
    extern __shared__ double lds[];
    constexpr int NS = PASS::NSLOT, NO = PASS::NOUT, NE = PASS::NEARLY;
    constexpr bool RE = PASS::REUSE;
    static_assert(NO <= NS, "late outputs must fit in the input slots");
    static_assert(NO > 0 || RE, "a pass without late outputs stores through store_early");
    static_assert(!RE || NE <= NS, "early outputs must fit in the input slots");
    constexpr int NPF = PASS::NPF;
    static_assert(NPF == 0 || PASS::LOAD_UNROLL >= M, "prefetching needs the load loop unrolled");
    double pf[M][NPF > 0 ? NPF : 1];
    const Walk<M, P, NT> W(g, tile_of(g));
    constexpr int ss = NT * M;           // slot size: B*M*P with B*P = NT

    // 1. load (memory order, coalesced)
#pragma unroll(PASS::LOAD_UNROLL)
    for (int r = 0; r < M; r++) {
        int b, s;
        const long gi = W.at(g, r, b, s);
        double v[NS > 0 ? NS : 1];
        PASS::load(g, a, gi, v);
        if constexpr (NPF > 0) PASS::prefetch(g, a, gi, pf[r]);
        const int k = s % M, c = s / M;
        const int li = lds_idx<M>(b, k, c, P);
#pragma unroll
        for (int q = 0; q < NS; q++) lds[q * ss + li] = v[q];
    }
    __syncthreads();

    // 2. compute (line mapping)
    const Lane<M, P> L = make_lane<M, P>(lds, ss);
    double out[NO > 0 ? NO : 1][M];
    PASS::template compute<M, P>(g, a, L, out);

    // 3. store (through LDS, memory order)
    __syncthreads();
    if (RE) {
#pragma unroll
        for (int r = 0; r < M; r++) {
            int b, s;
            const long gi = W.at(g, r, b, s);
            const int li = lds_idx<M>(b, s % M, s / M, P);
            double v[NE + NPF > 0 ? NE + NPF : 1];
#pragma unroll
            for (int j = 0; j < NE; j++) v[j] = lds[j * ss + li];
#pragma unroll
            for (int j = 0; j < NPF; j++) v[NE + j] = pf[r][j];
            PASS::store_early(g, a, gi, v);
        }
        if constexpr (NO == 0) return;
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NO; j++) {
#pragma unroll
        for (int k = 0; k < M; k++) L.st(j, k, out[j][k]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < M; r++) {
        int b, s;
        const long gi = W.at(g, r, b, s);
        const int k = s % M, c = s / M;
        const int li = lds_idx<M>(b, k, c, P);
        constexpr int NV = RE ? NO : NO + NE;
        double v[NV + NPF > 0 ? NV + NPF : 1];
#pragma unroll
        for (int j = 0; j < NO; j++) v[j] = lds[j * ss + li];
        if (!RE) {
#pragma unroll
            for (int j = NO; j < NV; j++) v[j] = lds[(NS + j - NO) * ss + li];
        }
#pragma unroll
        for (int j = 0; j < NPF; j++) v[NV + j] = pf[r][j];
        PASS::store(g, a, gi, v);
    }
}

// Passes whose early outputs have their own slots.
struct NoReuse {
    static constexpr bool REUSE = false;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void store_early(const Geo &, const PassArgs &, long, const double *) {}
};

// ---------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------
DEV double qf(const PassArgs &a, int comp, long gi) { return a.Q[comp * a.N + gi]; }

template <int M, int P>
DEV void qown(const Lane<M, P> &L, int k, double &q1, double &q2, double &q3, double &q4, double &q5)
{
    q1 = L.ld(0, k);
    q2 = L.ld(1, k);
    q3 = L.ld(2, k);
    q4 = L.ld(3, k);
    q5 = L.ld(4, k);
}

DEV void upd(double *p, double v, bool first)
{
    *p = first ? v : *p + v;
}

// ---------------------------------------------------------------------
// single-operator passes (the DFDX/D2FDX2/Compact_Filter_X/Gaussian_X
// entry points of the C-ABI)
// ---------------------------------------------------------------------
struct PassD1 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.in0[gi]; }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        a.out0[gi] = a.iadd ? a.out0[gi] + v[0] : v[0];
    }
};

struct PassD2 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 1;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.in0[gi]; }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO], d2[M];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, d2);
#pragma unroll
        for (int k = 0; k < M; k++) L.st(1, k, d2[k]);
        if (a.iadd) op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, out[0]);
        else {
#pragma unroll
            for (int k = 0; k < M; k++) out[0][k] = 0.0;
        }
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        if (a.iadd) a.out0[gi] = v[0];
        a.out1[gi] = v[1];
    }
};

struct PassFilt : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.in0[gi]; }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_filt<M>(w, a.C, a.tab->op[2].scale, a.tab->op[2], L, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v) { a.out0[gi] = v[0]; }
};

struct PassGauss : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.in0[gi]; }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_gauss<M>(w, a.C, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v) { a.out0[gi] = v[0]; }
};

// ---------------------------------------------------------------------
// P1: first derivatives, own-direction second derivatives and convective
// terms of one direction, from the state Q.
//   RHS_x  src/rhsv_periodic_sf.f:36-461
//   RHS_yz src/rhsv_periodic_sf.f:546-880 (y and z parts of the same terms)
// The load phase stages the primitives (u, v, w, T, rho) of
// src/rhsv_periodic_sf.f:92-104.  Per direction d (velocity u_d):
//   dQ_i += D1(-rho u_i u_d/2 - delta_id P) - D1(rho u_d) u_i/2
//           - (rho u_d/2) du_i/dx_d + w_id mu d2u_i/dx_d^2,  w_id = 4/3+bulk (i=d) else 1
//   dQ_4 += -D1(rho u_d)
//   dQ_5 += D1(-(rho T + rho |u|^2/2) u_d) + sum_i u_i w_id mu d2u_i/dx_d^2
//           + Pr^-1 (mu d2T/dx_d^2 + mu' (dT/dx_d)^2)
//   G(., d) = D1(u, v, w, T);  L2e += D2(T)/gamma   (e = T/gamma)
// D1 of a sum replaces the reference's sum of D1s of the same operator,
// rho*E + P = rho T + rho |u|^2 / 2, and D2(T)/gamma replaces D2(T/gamma);
// all three differ from the reference's evaluation only by rounding.
// ---------------------------------------------------------------------
// MW: occupancy target (waves per SIMD) handed to the register allocator;
// 3 suits the unit-stride y lines, the strided x and z lines run faster with
// the compiler's own choice.
template <int MW>
struct PassGrad {
    static constexpr int NSLOT = 5, NOUT = 5, NEARLY = 5;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = MW;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[5])
    {
        const Prim pr = prim(a.ph, a.Q[gi], a.Q[a.N + gi], a.Q[2 * a.N + gi], a.Q[3 * a.N + gi], a.Q[4 * a.N + gi]);
        v[0] = pr.u;
        v[1] = pr.v;
        v[2] = pr.w;
        v[3] = pr.T;
        v[4] = a.Q[3 * a.N + gi];
    }
    // Accumulate into row i of out (one of the rows in MASK) with static
    // register indices.
    template <int MASK, int M>
    DEV static void acc(double (&out)[5][M], int i, int k, double v)
    {
#pragma unroll
        for (int r = 0; r < 5; r++)
            if ((MASK >> r) & 1) out[r][k] += (i == r) ? v : 0.0;
    }
    // Operator order: fluxes and the mass flux first, while u, v, w, T and
    // rho are all needed; then T, whose slot takes dT/dx_d and rho's slot
    // the Laplacian of e; then each velocity, whose slot takes its
    // derivative.  The loop bodies are free of branches so that the
    // accumulators stay in place across iterations.
    template <int M, int P>
    DEV static void compute(const Geo &g, const PassArgs &a, const Lane<M, P> &L, double (&out)[5][M])
    {
        const Phys &ph = a.ph;
        const int d = g.dir;
        const double pfac = ph.gm1 * ph.ginv;
        double mu[M], rud[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            mu[k] = visc(ph, L.ld(3, k));
            rud[k] = -0.5 * L.ld(4, k) * L.ld(d, k);
#pragma unroll
            for (int i = 0; i < 5; i++) out[i][k] = 0.0;
        }
        // flux derivatives: momentum (with pressure) 0..2, energy 4
#pragma unroll 1
        for (int it = 0; it < 4; it++) {
            const int comp = it == 3 ? 4 : it;
            const int ci = it == 3 ? 0 : it;
            const double pw = comp == d ? 1.0 : 0.0;
            double f[M], w[M + 2 * HALO], df[M];
#pragma unroll
            for (int k = 0; k < M; k++) {
                const double u = L.ld(0, k), v = L.ld(1, k), ww = L.ld(2, k), T = L.ld(3, k), rho = L.ld(4, k);
                const double ud = L.ld(d, k), ui = L.ld(ci, k);
                const double fm = -0.5 * rho * ui * ud - pw * (rho * T * pfac);
                const double fe = -((rho * T + 0.5 * rho * (u * u + v * v + ww * ww)) * ud);
                f[k] = it == 3 ? fe : fm;
            }
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, df);
#pragma unroll
            for (int k = 0; k < M; k++) acc<0x17, M>(out, comp, k, df[k]);
        }
        // skew-symmetric part and continuity with the mass flux rho u_d
        {
            double f[M], w[M + 2 * HALO], sk[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(4, k) * L.ld(d, k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, sk);
#pragma unroll
            for (int k = 0; k < M; k++) {
                const double t = -0.5 * sk[k];
                out[0][k] += t * L.ld(0, k);
                out[1][k] += t * L.ld(1, k);
                out[2][k] += t * L.ld(2, k);
                out[3][k] = -sk[k];
            }
        }
        // T: first and second derivatives; the Laplacian of e = T/gamma is
        // D2(T)/gamma (the operator is linear)
        {
            double f[M], w[M + 2 * HALO], gd[M], hd[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(3, k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, gd);
            op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, hd);
#pragma unroll
            for (int k = 0; k < M; k++) {
                L.st(3, k, gd[k]);
                L.st(4, k, hd[k] * ph.ginv);
                const double mup = dvisc(ph, f[k], mu[k]);
                out[4][k] += ph.pr_inv * (mu[k] * hd[k] + mup * gd[k] * gd[k]);
            }
        }
        // u, v, w: first and second derivatives
#pragma unroll 1
        for (int comp = 0; comp < 3; comp++) {
            double f[M], w[M + 2 * HALO], gd[M], hd[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(comp, k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, gd);
            op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, hd);
            const double wd = comp == d ? (4.0 / 3.0 + ph.bulk) : 1.0;
#pragma unroll
            for (int k = 0; k < M; k++) {
                L.st(comp, k, gd[k]);
                const double vt = wd * mu[k] * hd[k];
                acc<0x07, M>(out, comp, k, vt + rud[k] * gd[k]);
                out[4][k] += f[k] * vt;
            }
        }
    }
    DEV static void store_early(const Geo &g, const PassArgs &a, long gi, const double *v)
    {
#pragma unroll
        for (int c = 0; c < 3; c++) a.G[(g.dir * 3 + c) * a.N + gi] = v[c];
        a.G[(9 + g.dir) * a.N + gi] = v[3];
        if (a.L2e) upd(&a.L2e[gi], v[4], a.first);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v)
    {
#pragma unroll
        for (int c = 0; c < 5; c++) upd(&a.dQ[c * a.N + gi], v[c], a.first);
    }
};

// ---------------------------------------------------------------------
// P2: mixed second derivatives of the viscous terms (D1 of stored first
// derivatives, src/rhsv_periodic_sf.f:1062-1070,1281-1363) and, with the
// artificial transport on, the second derivatives of |S| and of the
// internal energy's Laplacian (src/hyper_dynamic.f:272-276).
//   y pass: D1y(dux/dx), D1y(duy/dx), D1y(duy/dz), D1y(duz/dz)
//   z pass: D1z(dux/dx), D1z(duz/dx)
// ---------------------------------------------------------------------
template <int DIR, bool HYP>
struct PassMix {
    static constexpr int NMIX = DIR == 1 ? 4 : DIR == 2 ? 2 : 0;
    static constexpr int NH = HYP ? 2 : 0;
    static constexpr int NSLOT = NMIX + NH, NOUT = 0, NEARLY = NMIX + NH;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    // prefetched for the store phase: (1/3 + bulk) mu, u, v, w and the
    // right-hand sides 1, 3, 5 (and 2 for y); L2S and L4e when they are
    // accumulated (not the first mix pass)
    static constexpr int NPQ = NMIX > 0 ? 4 + (DIR == 1 ? 4 : 3) : 0;
    static constexpr int NPF = NPQ + (HYP && DIR != 0 ? 2 : 0);
    DEV static void prefetch(const Geo &, const PassArgs &a, long gi, double *p)
    {
        if (NMIX > 0) {
            const Phys &ph = a.ph;
            const Prim pr = prim(ph, qf(a, 0, gi), qf(a, 1, gi), qf(a, 2, gi), qf(a, 3, gi), qf(a, 4, gi));
            p[0] = (ph.bulk + 1.0 / 3.0) * visc(ph, pr.T);
            p[1] = pr.u;
            p[2] = pr.v;
            p[3] = pr.w;
            p[4] = a.dQ[gi];
            p[5] = a.dQ[2 * a.N + gi];
            p[6] = a.dQ[4 * a.N + gi];
            if (DIR == 1) p[7] = a.dQ[a.N + gi];
        }
        if (HYP && DIR != 0) {
            p[NPQ] = a.L2S[gi];
            p[NPQ + 1] = a.L4e[gi];
        }
    }
    // G index of each mixed input: gxx=0, gyx=1, gzx=2, gxy=3, gyy=4, gzy=5, gxz=6, gyz=7, gzz=8
    DEV static int gsrc(int m)
    {
        if (DIR == 1) return m == 0 ? 0 : m == 1 ? 1 : m == 2 ? 7 : 8;
        return m == 0 ? 0 : 2;
    }
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[NSLOT])
    {
#pragma unroll
        for (int m = 0; m < NMIX; m++) v[m] = a.G[gsrc(m) * a.N + gi];
        if (HYP) {
            v[NMIX] = a.S[gi];
            v[NMIX + 1] = a.L2e[gi];
        }
    }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&)[1][M])
    {
        // each result replaces its own input in LDS
#pragma unroll 1
        for (int m = 0; m < NMIX; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
#pragma unroll 1
        for (int m = NMIX; m < NSLOT; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        const double *p = v + NSLOT;
        if (NMIX > 0) {
            const double c3 = p[0];
            double fx, fy, fz;
            if (DIR == 1) {       // v: d/dy(dux/dx), d/dy(duy/dx), d/dy(duy/dz), d/dy(duz/dz)
                fx = c3 * v[1];
                fy = c3 * (v[0] + v[3]);
                fz = c3 * v[2];
            } else {              // v: d/dz(dux/dx), d/dz(duz/dx)
                fx = c3 * v[1];
                fy = 0.0;
                fz = c3 * v[0];
            }
            a.dQ[gi] = p[4] + fx;
            if (DIR == 1) a.dQ[a.N + gi] = p[7] + fy;
            a.dQ[2 * a.N + gi] = p[5] + fz;
            a.dQ[4 * a.N + gi] = p[6] + (p[1] * fx + p[2] * fy + p[3] * fz);
        }
        if (HYP) {
            if (DIR == 0) {
                upd(&a.L2S[gi], v[NMIX], a.first);
                upd(&a.L4e[gi], v[NMIX + 1], a.first);
            } else {
                a.L2S[gi] = a.first ? v[NMIX] : p[NPQ] + v[NMIX];
                a.L4e[gi] = a.first ? v[NMIX + 1] : p[NPQ + 1] + v[NMIX + 1];
            }
        }
    }
};

// ---------------------------------------------------------------------
// P3: Laplacian of Lap|S|.  The z pass (last) also forms the transport
// drivers eta = (rho |Lap^2 S|, (rho/T) c |Lap^2 e|)
// (src/hyper_dynamic.f:278-297).
// ---------------------------------------------------------------------
template <bool LAST>
struct PassLap2 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.L2S[gi]; }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        if (!LAST) {
            upd(&a.L4S[gi], v[0], a.first);
            return;
        }
        const Phys &ph = a.ph;
        const double d4S = fabs(a.L4S[gi] + v[0]);
        const double d4e = fabs(a.L4e[gi]);
        const double q4 = qf(a, 3, gi);
        const Prim pr = prim(ph, qf(a, 0, gi), qf(a, 1, gi), qf(a, 2, gi), q4, qf(a, 4, gi));
        const double fac3 = (q4 / pr.T) * sqrt(ph.gamma * pr.p * pr.vs);
        a.eta[gi] = q4 * d4S;
        a.eta[a.N + gi] = fac3 * d4e;
    }
};

// ---------------------------------------------------------------------
// Gaussian smoothing of eta (in place).  The z pass (last) forms the
// transport coefficients beta*, mu*, kappa* (src/hyper_dynamic.f:303-314).
// ---------------------------------------------------------------------
template <bool LAST>
struct PassGaussEta : NoReuse {
    static constexpr int NSLOT = 2, NOUT = 2, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[2])
    {
        v[0] = a.eta[gi];
        v[1] = a.eta[a.N + gi];
    }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&out)[2][M])
    {
#pragma unroll
        for (int m = 0; m < 2; m++) {
            double f[M], w[M + 2 * HALO];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_gauss<M>(w, a.C, out[m]);
        }
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        if (!LAST) {
            a.eta[gi] = v[0];
            a.eta[a.N + gi] = v[1];
            return;
        }
        const Phys &ph = a.ph;
        a.Hc[gi] = a.Cb[gi] * v[0] * ph.dx6;            // beta*
        a.Hc[a.N + gi] = ph.cmu * v[0] * ph.dx6;        // mu*
        a.Hc[2 * a.N + gi] = ph.ck * v[1] * ph.dx5;     // kappa*
    }
};

// ---------------------------------------------------------------------
// P5: divergence of the artificial stress and heat flux
// (src/hyper_rhs.f:159-491).  For direction d the staged fields are the
// momentum fluxes tau_xd, tau_yd, tau_zd and the energy flux
// u tau_xd + v tau_yd + w tau_zd + kappa* dT/dx_d (the reference takes the
// D1 of each energy term separately; the sum of the D1s is the D1 of the
// sum).
// ---------------------------------------------------------------------
struct PassHyperDiv {
    static constexpr int NSLOT = 4, NOUT = 0, NEARLY = 4;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 4;            // dQ 1, 2, 3, 5 before this pass
    DEV static void prefetch(const Geo &, const PassArgs &a, long gi, double *p)
    {
        p[0] = a.dQ[gi];
        p[1] = a.dQ[a.N + gi];
        p[2] = a.dQ[2 * a.N + gi];
        p[3] = a.dQ[4 * a.N + gi];
    }
    DEV static void load(const Geo &g, const PassArgs &a, long gi, double (&v)[4])
    {
        const long N = a.N;
        const double *G = a.G;
        const double beta = a.Hc[gi], mu = a.Hc[N + gi], kap = a.Hc[2 * N + gi];
        const double gxx = G[gi], gyy = G[4 * N + gi], gzz = G[8 * N + gi];
        const double dil = gxx + gyy + gzz;
        const double q4 = a.Q[3 * N + gi];
        const double uh = a.Q[gi] / q4, vh = a.Q[N + gi] / q4, wh = a.Q[2 * N + gi] / q4;
        const double tt = 2.0 / 3.0;
        const double bulkterm = (beta - tt * mu) * dil;
        double tx, ty, tz;
        if (g.dir == 0) {
            tx = 2.0 * mu * gxx + bulkterm;
            ty = mu * (G[3 * N + gi] + G[1 * N + gi]);
            tz = mu * (G[6 * N + gi] + G[2 * N + gi]);
        } else if (g.dir == 1) {
            tx = mu * (G[3 * N + gi] + G[1 * N + gi]);
            ty = 2.0 * mu * gyy + bulkterm;
            tz = mu * (G[7 * N + gi] + G[5 * N + gi]);
        } else {
            tx = mu * (G[6 * N + gi] + G[2 * N + gi]);
            ty = mu * (G[7 * N + gi] + G[5 * N + gi]);
            tz = 2.0 * mu * gzz + bulkterm;
        }
        v[0] = tx;
        v[1] = ty;
        v[2] = tz;
        v[3] = uh * tx + vh * ty + wh * tz + kap * G[(9 + g.dir) * N + gi];
    }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&)[1][M])
    {
#pragma unroll 1
        for (int m = 0; m < 4; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        a.dQ[gi] = v[4] + v[0];
        a.dQ[a.N + gi] = v[5] + v[1];
        a.dQ[2 * a.N + gi] = v[6] + v[2];
        a.dQ[4 * a.N + gi] = v[7] + v[3];
    }
};

// ---------------------------------------------------------------------
// Low-storage RK update fused into the x pass of the compact filter
// (src/lddrk4_periodic_sf.f:349-354), then the y and z filter passes; the
// z pass applies Fix_Pressure (src/lddrk4_periodic_sf.f:436-473).
// ---------------------------------------------------------------------
template <int KIND>   // 0: RK update + x filter, 1: y filter, 2: z filter + pressure fix
struct PassFilterQ {
    static constexpr int NSLOT = 5, NOUT = 0, NEARLY = 5;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = 1;
    // the RK update loads 15 fields per point: two points in flight at once
    static constexpr int LOAD_UNROLL = KIND == 0 ? 2 : 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[5])
    {
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const long o = c * a.N + gi;
            if (KIND == 0) {
                const double qw = a.rk_a * a.QW[o] + a.dt * a.dQ[o];
                a.QW[o] = qw;
                v[c] = a.Q[o] + a.rk_b * qw;
            } else {
                v[c] = a.Q[o];
            }
        }
    }
    template <int M, int P>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P> &L, double (&)[1][M])
    {
        const OpTab &T = a.tab->op[2];
#pragma unroll 1
        for (int m = 0; m < 5; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_filt<M>(w, a.C, T.scale, T, L, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        const long N = a.N;
        if (KIND == 2 && a.fix) {
            const Phys &ph = a.ph;
            const double vs = 1.0 / v[3];
            const double ke = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            double p = ph.gm1 * (v[4] - 0.5 * vs * ke);
            if (p < 0.0) p = 0.71463;
            a.Q[gi] = v[0];
            a.Q[N + gi] = v[1];
            a.Q[2 * N + gi] = v[2];
            a.Q[3 * N + gi] = v[3];
            a.Q[4 * N + gi] = ph.gm1_inv * p + 0.5 * vs * ke;
            return;
        }
#pragma unroll
        for (int c = 0; c < 5; c++) a.Q[c * N + gi] = v[c];
    }
};

// ---------------------------------------------------------------------
// pointwise kernels
// ---------------------------------------------------------------------

// Viscous terms built from products of first derivatives
// (src/rhsv_periodic_sf.f:1128-1154 without the mixed and own-direction
// second derivatives, and :1252-1275), plus |S| (src/hyper_dynamic.f:12-33,
// 227-245) and the dilatation sensor C_beta (src/hyper_dynamic.f:142-160).
__global__ void __launch_bounds__(256) k_cross(PassArgs a, int hyp)
{
    const long N = a.N;
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < N; gi += (long)gridDim.x * blockDim.x) {
        const Phys &ph = a.ph;
        const double q1 = a.Q[gi], q2 = a.Q[N + gi], q3 = a.Q[2 * N + gi], q4 = a.Q[3 * N + gi], q5 = a.Q[4 * N + gi];
        const Prim pr = prim(ph, q1, q2, q3, q4, q5);
        const double *G = a.G;
        const double gxx = G[gi], gyx = G[N + gi], gzx = G[2 * N + gi];
        const double gxy = G[3 * N + gi], gyy = G[4 * N + gi], gzy = G[5 * N + gi];
        const double gxz = G[6 * N + gi], gyz = G[7 * N + gi], gzz = G[8 * N + gi];
        const double mu = visc(ph, pr.T), mup = dvisc(ph, pr.T, mu);
        const double tx = mup * G[9 * N + gi], ty = mup * G[10 * N + gi], tz = mup * G[11 * N + gi];
        const double f1 = ph.bulk - 2.0 / 3.0;
        const double dil = gxx + gyy + gzz;
        const double t1 = gyx + gxy, t2 = gzx + gxz, t3 = gzy + gyz;
        const double fX = tx * (2.0 * gxx + f1 * dil) + ty * t1 + tz * t2;
        const double fY = ty * (2.0 * gyy + f1 * dil) + tx * t1 + tz * t3;
        const double fZ = tz * (2.0 * gzz + f1 * dil) + tx * t2 + ty * t3;
        // the y pass left its terms in dQ, the z and x passes theirs in dQp
        const double *Pz = a.dQp, *Px = a.dQp + 5 * N;
        a.dQ[gi] = ((a.dQ[gi] + Pz[gi]) + Px[gi]) + fX;
        a.dQ[N + gi] = ((a.dQ[N + gi] + Pz[N + gi]) + Px[N + gi]) + fY;
        a.dQ[2 * N + gi] = ((a.dQ[2 * N + gi] + Pz[2 * N + gi]) + Px[2 * N + gi]) + fZ;
        a.dQ[3 * N + gi] = (a.dQ[3 * N + gi] + Pz[3 * N + gi]) + Px[3 * N + gi];
        a.dQ[4 * N + gi] = ((a.dQ[4 * N + gi] + Pz[4 * N + gi]) + Px[4 * N + gi])
                           + (pr.u * fX + pr.v * fY + pr.w * fZ
                              + mu * (2.0 * (gxx * gxx + gyy * gyy + gzz * gzz) + f1 * dil * dil)
                              + mu * (t1 * t1 + t2 * t2 + t3 * t3));
        if (a.L2e) a.L2e[gi] = (a.L2e[gi] + a.L2ep[gi]) + a.L2ep[N + gi];
        if (hyp) {
            const double s12 = 0.5 * (gxy + gyx), s13 = 0.5 * (gxz + gzx), s23 = 0.5 * (gyz + gzy);
            a.S[gi] = sqrt(0.5 * (gxx * gxx + gyy * gyy + gzz * gzz + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
            const double wx = gzy - gyz, wy = gxz - gzx, wz = gyx - gxy;
            const double om2 = wx * wx + wy * wy + wz * wz;
            const double fac = 0.5 * (1.0 - tanh(2.5 + 10.0 * dil));
            a.Cb[gi] = fac * dil * dil / (dil * dil + om2 + 1.0e-7) + 1.0e-20;
        }
    }
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
template <int M, int P, int NT, class PASS>
static hipError_t launch_t(const Geo &g, const PassArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)pass_slots<PASS>() * NT * M * sizeof(double);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    static size_t lds_set = 0;
    if (lds > 64 * 1024 && lds > lds_set) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pass<M, P, NT, PASS>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    hipLaunchKernelGGL((k_pass<M, P, NT, PASS>), dim3((unsigned)g.ntiles), dim3(NT), lds, s, g, a);
    return hipGetLastError();
}

// The lane layouts (M points per lane, P lanes per line) that
// choose_layout and its narrow variant (M/2, 2P) produce for lines of
// 16..512 points.
template <class PASS, int NT>
static hipError_t launch_m(int M, int P, const Geo &g, const PassArgs &a, hipStream_t s)
{
    switch (M * 100 + P) {
    case 116: return launch_t<1, 16, NT, PASS>(g, a, s);
    case 208: return launch_t<2, 8, NT, PASS>(g, a, s);
    case 216: return launch_t<2, 16, NT, PASS>(g, a, s);
    case 408: return launch_t<4, 8, NT, PASS>(g, a, s);
    case 416: return launch_t<4, 16, NT, PASS>(g, a, s);
    case 432: return launch_t<4, 32, NT, PASS>(g, a, s);
    case 464: return launch_t<4, 64, NT, PASS>(g, a, s);
    case 808: return launch_t<8, 8, NT, PASS>(g, a, s);
    case 816: return launch_t<8, 16, NT, PASS>(g, a, s);
    case 832: return launch_t<8, 32, NT, PASS>(g, a, s);
    case 864: return launch_t<8, 64, NT, PASS>(g, a, s);
    default: return hipErrorInvalidValue;
    }
}

// t holds the tables of the chosen layout (M points per lane, P lanes per
// line); nt threads per workgroup give B = nt/P lines per tile.
template <class PASS>
static hipError_t launch(const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    Geo g;
    g.dir = dir;
    g.n = t.n;
    g.My = My;
    g.Mz = Mz;
    g.Mx = Mx;
    g.npts = (long)My * Mz * Mx;
    g.P = t.P;
    g.B = nt / t.P;
    if (g.B < 1 || g.B * t.P != nt) return hipErrorInvalidValue;
    const long lot = g.npts / g.n;
    if (lot % g.B) return hipErrorInvalidValue;
    g.ntiles = lot / g.B;
    if (nt == 256) return launch_m<PASS, 256>(t.M, t.P, g, a, s);
    if (nt == 512) return launch_m<PASS, 512>(t.M, t.P, g, a, s);
    return hipErrorInvalidValue;
}

void pass_default_cfg(int kind, int *narrow, int *nt)
{
    // measured per pass at 256^3: 512-thread tiles where the pass is bound
    // by LDS occupancy rather than by registers; M/2 points per lane
    // ("narrow") where the pass holds many fields in registers
    switch (kind) {
    case K_GRAD: case K_MIX_Y: case K_MIX_Z: case K_MIXH_Y: case K_MIXH_Z:
        *narrow = 1;
        *nt = 256;
        break;
    case K_MIXH_X: case K_GAUSS_ETA: case K_GAUSS_ETA_LAST:
    case K_HDIV: case K_RKF_X: case K_F_Y: case K_F_Z:
        *narrow = 1;
        *nt = 512;
        break;
    case K_LAP2_LAST:
        *narrow = 0;
        *nt = 512;
        break;
    default:
        *narrow = 0;
        *nt = 256;
    }
}

hipError_t run_pass(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_D1: return launch<PassD1>(t, nt, dir, My, Mz, Mx, a, s);
    case K_D2: return launch<PassD2>(t, nt, dir, My, Mz, Mx, a, s);
    case K_FILT: return launch<PassFilt>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS: return launch<PassGauss>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GRAD:
        if (dir == 1) return launch<PassGrad<3>>(t, nt, dir, My, Mz, Mx, a, s);
        return launch<PassGrad<1>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIX_Y: return launch<PassMix<1, false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIX_Z: return launch<PassMix<2, false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_X: return launch<PassMix<0, true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_Y: return launch<PassMix<1, true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_Z: return launch<PassMix<2, true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_LAP2: return launch<PassLap2<false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_LAP2_LAST: return launch<PassLap2<true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA: return launch<PassGaussEta<false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA_LAST: return launch<PassGaussEta<true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_HDIV: return launch<PassHyperDiv>(t, nt, dir, My, Mz, Mx, a, s);
    case K_RKF_X: return launch<PassFilterQ<0>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_F_Y: return launch<PassFilterQ<1>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_F_Z: return launch<PassFilterQ<2>>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t run_cross(const PassArgs &a, int hyp, hipStream_t s)
{
    const long N = a.N;
    long blocks = (N + 255) / 256;
    if (blocks > 2048 * 8) blocks = 2048 * 8;
    hipLaunchKernelGGL(k_cross, dim3((unsigned)blocks), dim3(256), 0, s, a, hyp);
    return hipGetLastError();
}

}  // namespace st

// ---------------------------------------------------------------------
// reductions (one partial per block; the host folds the partials)
// ---------------------------------------------------------------------
namespace st {

template <int NV, class OP>
DEV void block_reduce(double (&v)[NV], OP op, double *partial)
{
    __shared__ double sh[NV][256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double x = v[i];
        for (int o = 32; o > 0; o >>= 1) x = op(x, __shfl_xor(x, o));
        if (lane == 0) sh[i][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double x = sh[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) x = op(x, sh[threadIdx.x][w]);
        partial[blockIdx.x * NV + threadIdx.x] = x;
    }
}

struct OpMax { DEV double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpMin { DEV double operator()(double a, double b) const { return a < b ? a : b; } };
struct OpSum { DEV double operator()(double a, double b) const { return a + b; } };

// Compute_CFL src/lddrk4_periodic_sf.f:29-56
__global__ void __launch_bounds__(256) k_cfl(PassArgs a, double *partial)
{
    const long N = a.N;
    const Phys &ph = a.ph;
    const double fac = ph.gamma * ph.gm1;
    double v[2] = {-INFINITY, -INFINITY};
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < N; gi += (long)gridDim.x * blockDim.x) {
        const double q1 = a.Q[gi], q2 = a.Q[N + gi], q3 = a.Q[2 * N + gi], q4 = a.Q[3 * N + gi], q5 = a.Q[4 * N + gi];
        const double t1 = 1.0 / q4;
        const double t2 = sqrt(fac * t1 * (q5 - 0.5 * t1 * (q1 * q1 + q2 * q2 + q3 * q3)));
        const double ux = fabs(q1 * t1), uy = fabs(q2 * t1), uz = fabs(q3 * t1);
        const double w1 = (t2 + ux) * 2.478 / ph.D[0] + (t2 + uy) * 2.478 / ph.D[1] + (t2 + uz) * 2.478 / ph.D[2];
        const double w2 = (t2 + ux) * 2.199 / ph.D[0] + (t2 + uy) * 2.199 / ph.D[1] + (t2 + uz) * 2.199 / ph.D[2];
        v[0] = w1 > v[0] ? w1 : v[0];
        v[1] = w2 > v[1] ? w2 : v[1];
    }
    block_reduce<2>(v, OpMax(), partial);
}

// Compute_Tstep src/lddrk4_periodic_sf.f:105-132 (beta*, mu*, kappa* in Hc)
__global__ void __launch_bounds__(256) k_tstep(PassArgs a, double *partial)
{
    const long N = a.N;
    const Phys &ph = a.ph;
    const double fac = ph.gamma * ph.gm1;
    const double del = ph.D[0];
    double v[3] = {INFINITY, INFINITY, INFINITY};
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < N; gi += (long)gridDim.x * blockDim.x) {
        const double q1 = a.Q[gi], q2 = a.Q[N + gi], q3 = a.Q[2 * N + gi], q4 = a.Q[3 * N + gi], q5 = a.Q[4 * N + gi];
        const double t1 = 1.0 / q4;
        const double br = q5 - 0.5 * t1 * (q1 * q1 + q2 * q2 + q3 * q3);
        const double t2 = sqrt(fac * t1 * br);
        const double T = ph.gamma * t1 * br;
        const double w3 = q4 * (del * del) / a.Hc[N + gi];
        const double w4 = q4 * (del * del) / a.Hc[gi];
        const double w5 = q4 * ((del * t2) * (del * t2)) / (a.Hc[2 * N + gi] * T);
        v[0] = w3 < v[0] ? w3 : v[0];
        v[1] = w4 < v[1] ? w4 : v[1];
        v[2] = w5 < v[2] ? w5 : v[2];
    }
    block_reduce<3>(v, OpMin(), partial);
}

// Compute_Mom/KE/Mass/Energy/Entropy src/lddrk4_periodic_sf.f:153-302 (sums)
__global__ void __launch_bounds__(256) k_diag(PassArgs a, double *partial)
{
    const long N = a.N;
    const Phys &ph = a.ph;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < N; gi += (long)gridDim.x * blockDim.x) {
        const double q1 = a.Q[gi], q2 = a.Q[N + gi], q3 = a.Q[2 * N + gi], q4 = a.Q[3 * N + gi], q5 = a.Q[4 * N + gi];
        const double t = 1.0 / q4;
        const double u1 = q1 * t, u2 = q2 * t, u3 = q3 * t;
        v[0] += q1;
        v[1] += q2;
        v[2] += q3;
        v[3] += q4 * u1 * u1 + q4 * u2 * u2 + q4 * u3 * u3;
        v[4] += q4;
        v[5] += q5;
        const double p = ph.gm1 * (q5 - 0.5 * t * (q1 * q1 + q2 * q2 + q3 * q3));
        v[6] += q4 * log(p * pow(t, ph.gamma));
    }
    block_reduce<7>(v, OpSum(), partial);
}

__global__ void __launch_bounds__(256) k_fill(double *p, long n, double v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Fix_Pressure on a resident array (src/lddrk4_periodic_sf.f:436-473)
__global__ void __launch_bounds__(256) k_fixp(PassArgs a, double *Qx)
{
    const long N = a.N;
    const Phys &ph = a.ph;
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < N; gi += (long)gridDim.x * blockDim.x) {
        const double q1 = Qx[gi], q2 = Qx[N + gi], q3 = Qx[2 * N + gi], q4 = Qx[3 * N + gi], q5 = Qx[4 * N + gi];
        const double vs = 1.0 / q4;
        const double ke = q1 * q1 + q2 * q2 + q3 * q3;
        double p = ph.gm1 * (q5 - 0.5 * vs * ke);
        if (p < 0.0) p = 0.71463;
        Qx[4 * N + gi] = ph.gm1_inv * p + 0.5 * vs * ke;
    }
}

static long red_blocks(long N)
{
    long b = (N + 255) / 256;
    return b > 1024 ? 1024 : b;
}

hipError_t run_reduce(int which, const PassArgs &a, double *partial, int *nblocks, hipStream_t s)
{
    const long nb = red_blocks(a.N);
    *nblocks = (int)nb;
    if (which == 0) hipLaunchKernelGGL(k_cfl, dim3((unsigned)nb), dim3(256), 0, s, a, partial);
    else if (which == 1) hipLaunchKernelGGL(k_tstep, dim3((unsigned)nb), dim3(256), 0, s, a, partial);
    else hipLaunchKernelGGL(k_diag, dim3((unsigned)nb), dim3(256), 0, s, a, partial);
    return hipGetLastError();
}

hipError_t run_fill(double *p, long n, double v, hipStream_t s)
{
    long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t run_fixp(const PassArgs &a, double *Qx, hipStream_t s)
{
    const long nb = red_blocks(a.N) * 4;
    hipLaunchKernelGGL(k_fixp, dim3((unsigned)nb), dim3(256), 0, s, a, Qx);
    return hipGetLastError();
}

}  // namespace st
