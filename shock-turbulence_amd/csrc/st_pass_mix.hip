// Mixed-derivative passes of the right-hand side.
#include "st_pass.hpp"

namespace st {

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
    template <int M, int P, bool NP>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP> &L, double (&)[1][M])
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


hipError_t launch_mix(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_MIX_Y: return launch<PassMix<1, false>, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIX_Z: return launch<PassMix<2, false>, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_X: return launch<PassMix<0, true>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_Y: return launch<PassMix<1, true>, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_MIXH_Z: return launch<PassMix<2, true>, false>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st
