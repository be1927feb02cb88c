// Divergence of the artificial stress and heat flux, shared by the
// hyper_div pass and its fused form with the stage update and x filter.
#pragma once

#include "st_pass.hpp"

namespace st {

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
        fluxes(g, a, gi, a.Hc[gi], a.Hc[N + gi], a.Hc[2 * N + gi], v);
    }
    // the staged fluxes of direction g.dir from the coefficients at the point
    // (SEL: the direction's gradient entries chosen by index, so that the
    // loads of a point issue together, which the load phases of hyper_div
    // and hdiv_rkf gain from (2-3%); the z pass of gauss_hdiv computes them
    // in its mid phase, where the branch on the direction measured faster)
    template <bool SEL = true>
    DEV static void fluxes(const Geo &g, const PassArgs &a, long gi, double beta, double mu, double kap,
                           double (&v)[4])
    {
        const long N = a.N;
        const double *G = a.G;
        const double gxx = G[gi], gyy = G[4 * N + gi], gzz = G[8 * N + gi];
        const double dil = gxx + gyy + gzz;
        const double q4 = a.Q[3 * N + gi];
        const double uh = a.Q[gi] / q4, vh = a.Q[N + gi] / q4, wh = a.Q[2 * N + gi] / q4;
        const double tt = 2.0 / 3.0;
        const double bulkterm = (beta - tt * mu) * dil;
        double tx, ty, tz;
        if constexpr (SEL) {
            // the two off-diagonal stresses of direction d from the pairs
            // (du_i/dx_j, du_j/dx_i) with the other two directions: x-y
            // (3, 1), x-z (6, 2), y-z (7, 5) (d is the same on every lane)
            const int d = g.dir;
            const int a1 = d == 2 ? 6 : 3, a2 = d == 2 ? 2 : 1;
            const int b1 = d == 0 ? 6 : 7, b2 = d == 0 ? 2 : 5;
            const double sa = mu * (G[a1 * N + gi] + G[a2 * N + gi]);
            const double sb = mu * (G[b1 * N + gi] + G[b2 * N + gi]);
            const double gdd = d == 0 ? gxx : d == 1 ? gyy : gzz;
            const double diag = 2.0 * mu * gdd + bulkterm;
            tx = d == 0 ? diag : sa;
            ty = d == 1 ? diag : d == 0 ? sa : sb;
            tz = d == 2 ? diag : sb;
        } else if (g.dir == 0) {
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
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
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

}  // namespace st
