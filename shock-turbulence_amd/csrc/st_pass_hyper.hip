// Artificial-transport passes: second Laplacians, Gaussian smoothing, flux divergence.
#include "st_pass_hdiv.hpp"

namespace st {

// ---------------------------------------------------------------------
// P3: Laplacian of Lap|S| (src/hyper_dynamic.f:272-279), x pass: adds D2x
// to the z pass's partial in L4S.
// ---------------------------------------------------------------------
struct PassLap2 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.L2S[gi]; }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v) { upd(&a.L4S[gi], v[0], a.first); }
};

// D2 along the sweep of the first staged slot, in place.
template <int M, int P, bool NP, bool AL>
DEV void d2_slot(const PassArgs &a, const Lane<M, P, NP, AL> &L, int m)
{
    double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
    for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
    window<M>(f, L, w);
    op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, r);
#pragma unroll
    for (int k = 0; k < M; k++) L.st(m, k, r[k]);
}

// ---------------------------------------------------------------------
// The z pass that completes Lap|S| and Lap(Lap e) (the last direction of
// the D2 of |S| and Lap e, after x and y) and starts Lap(Lap|S|) on the
// same z lines:
//   compute   D2z of |S| and Lap e
//   mid       per point: L2S and L4e += those (written)
//   compute2  D2z of L2S
//   store     L4S = D2z L2S (the first direction of the second Laplacian)
// ---------------------------------------------------------------------
struct PassLapSELap2 {
    static constexpr int NSLOT = 2, NOUT = 0, NEARLY = 1;
    static constexpr bool REUSE = true;
    static constexpr bool MID = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[2])
    {
        v[0] = a.S[gi];
        v[1] = a.L2e[gi];
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
    {
#pragma unroll 1
        for (int m = 0; m < 2; m++) d2_slot<M, P, NP>(a, L, m);
    }
    DEV static void mid(const Geo &, const PassArgs &a, long gi, int, double (&v)[2])
    {
        const double l2s = a.L2S[gi] + v[0];       // as the accumulating D2 pass of |S| and Lap e
        a.L2S[gi] = l2s;
        a.L4e[gi] = a.L4e[gi] + v[1];
        v[0] = l2s;
        v[1] = 0.0;
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute2(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L)
    {
        d2_slot<M, P, NP>(a, L, 0);
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v) { a.L4S[gi] = v[0]; }
};

// ---------------------------------------------------------------------
// The y pass that completes Lap(Lap|S|) (after z and x), forms the transport
// drivers eta = (rho |Lap^2 S|, (rho/T) c |Lap^2 e|) (src/hyper_dynamic.f:
// 278-297) and smooths them along y (the first Gaussian direction):
//   compute   D2y of L2S
//   mid       per point: |L4S + D2y|, |L4e| and the state give eta
//   compute2  Gaussian along y of both eta fields
//   store     eta
// ---------------------------------------------------------------------
struct PassLap2Gauss {
    static constexpr int NSLOT = 2, NOUT = 0, NEARLY = 2;
    static constexpr bool REUSE = true;
    static constexpr bool MID = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[2])
    {
        v[0] = a.L2S[gi];
        v[1] = 0.0;
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
    {
        d2_slot<M, P, NP>(a, L, 0);
    }
    DEV static void mid(const Geo &, const PassArgs &a, long gi, int, double (&v)[2])
    {
        const Phys &ph = a.ph;
        const double d4S = fabs(a.L4S[gi] + v[0]);
        const double d4e = fabs(a.L4e[gi]);
        const double q4 = qf(a, 3, gi);
        const Prim pr = prim(ph, qf(a, 0, gi), qf(a, 1, gi), qf(a, 2, gi), q4, qf(a, 4, gi));
        const double fac3 = (q4 / pr.T) * sqrt(ph.gamma * pr.p * pr.vs);
        v[0] = q4 * d4S;
        v[1] = fac3 * d4e;
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute2(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L)
    {
#pragma unroll 1
        for (int m = 0; m < 2; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_gauss<M>(w, a.C, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v)
    {
        a.eta[gi] = v[0];
        a.eta[a.N + gi] = v[1];
    }
};

// ---------------------------------------------------------------------
// Gaussian smoothing of eta (in place), the x pass (y is done in
// PassLap2Gauss, z in PassGaussHdiv).
// ---------------------------------------------------------------------
struct PassGaussEta : NoReuse {
    static constexpr int NSLOT = 2, NOUT = 2, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[2])
    {
        v[0] = a.eta[gi];
        v[1] = a.eta[a.N + gi];
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[2][M])
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
        a.eta[gi] = v[0];
        a.eta[a.N + gi] = v[1];
    }
};

// ---------------------------------------------------------------------
// The z pass of the Gaussian smoothing fused with the z pass of the
// divergence: the smoothed eta gives the transport coefficients beta*,
// mu*, kappa* (src/hyper_dynamic.f:303-314), which are written out for the
// other divergence passes and, at the same point, give the z fluxes of the
// divergence (src/hyper_rhs.f:159-491), whose D1z is added into dQ.
//   compute   Gaussian along z of the two eta fields
//   mid       per point: beta*, mu*, kappa* (written), the staged z fluxes
//   compute2  D1z of the fluxes
// ---------------------------------------------------------------------
struct PassGaussHdiv {
    static constexpr int NSLOT = 4, NOUT = 0, NEARLY = 4;
    static constexpr bool REUSE = true;
    static constexpr bool MID = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = PassHyperDiv::NPF;
    DEV static void prefetch(const Geo &g, const PassArgs &a, long gi, double *p) { PassHyperDiv::prefetch(g, a, gi, p); }
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[4])
    {
        v[0] = a.eta[gi];
        v[1] = a.eta[a.N + gi];
        v[2] = 0.0;
        v[3] = 0.0;
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
    {
#pragma unroll
        for (int m = 0; m < 2; m++) {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(m, k);
            window<M>(f, L, w);
            op_gauss<M>(w, a.C, r);
#pragma unroll
            for (int k = 0; k < M; k++) L.st(m, k, r[k]);
        }
    }
    DEV static void mid(const Geo &g, const PassArgs &a, long gi, int, double (&v)[4])
    {
        const Phys &ph = a.ph;
        const long N = a.N;
        // beta*; constant C_beta = 1 without the sensor (src/hyper_routines.f:99-101,204)
        const double beta = (a.Cb ? a.Cb[gi] : 1.0) * v[0] * ph.dx6;
        const double mu = ph.cmu * v[0] * ph.dx6;
        const double kap = ph.ck * v[1] * ph.dx5;
        a.Hc[gi] = beta;
        a.Hc[N + gi] = mu;
        a.Hc[2 * N + gi] = kap;
        PassHyperDiv::fluxes<false>(g, a, gi, beta, mu, kap, v);
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute2(const Geo &g, const PassArgs &a, const Lane<M, P, NP, AL> &L)
    {
        double o[1][M];
        PassHyperDiv::compute<M, P, NP>(g, a, L, o);
    }
    DEV static void store(const Geo &, const PassArgs &, long, const double *) {}
    DEV static void store_early(const Geo &g, const PassArgs &a, long gi, const double *v)
    {
        PassHyperDiv::store_early(g, a, gi, v);
    }
};


hipError_t launch_hyper(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_LAP2: return launch<PassLap2>(t, nt, dir, My, Mz, Mx, a, s);
    case K_LAP2_LAST: return launch<PassLap2Gauss, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_LAPSE_LAP2: return launch<PassLapSELap2, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA: return launch<PassGaussEta>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA_LAST: return launch<PassGaussHdiv, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_HDIV: return launch<PassHyperDiv>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_hyper_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    switch (kind) {
    case K_LAP2: return launch_ragged<PassLap2>(sp, s);
    case K_LAP2_LAST: return launch_ragged<PassLap2Gauss>(sp, s);
    case K_LAPSE_LAP2: return launch_ragged<PassLapSELap2>(sp, s);
    case K_GAUSS_ETA: return launch_ragged<PassGaussEta>(sp, s);
    case K_GAUSS_ETA_LAST: return launch_ragged<PassGaussHdiv>(sp, s);
    case K_HDIV: return launch_ragged<PassHyperDiv>(sp, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st
