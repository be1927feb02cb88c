// Artificial-transport passes: second Laplacians, Gaussian smoothing, flux divergence.
#include "st_pass_hdiv.hpp"

namespace st {

// ---------------------------------------------------------------------
// P3: Laplacian of Lap|S|.  The z pass (last) also forms the transport
// drivers eta = (rho |Lap^2 S|, (rho/T) c |Lap^2 e|)
// (src/hyper_dynamic.f:278-297).
// ---------------------------------------------------------------------
template <bool LAST>
struct PassLap2 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.L2S[gi]; }
    template <int M, int P, bool NP>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP> &L, double (&out)[1][M])
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
    template <int M, int P, bool NP>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP> &L, double (&out)[2][M])
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
        // beta*; constant C_beta = 1 without the sensor (src/hyper_routines.f:99-101,204)
        a.Hc[gi] = (a.Cb ? a.Cb[gi] : 1.0) * v[0] * ph.dx6;
        a.Hc[a.N + gi] = ph.cmu * v[0] * ph.dx6;        // mu*
        a.Hc[2 * a.N + gi] = ph.ck * v[1] * ph.dx5;     // kappa*
    }
};

hipError_t launch_hyper(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_LAP2: return launch<PassLap2<false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_LAP2_LAST: return launch<PassLap2<true>, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA: return launch<PassGaussEta<false>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS_ETA_LAST: return launch<PassGaussEta<true>, false>(t, nt, dir, My, Mz, Mx, a, s);
    case K_HDIV: return launch<PassHyperDiv>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st
