// Second derivatives of |S| and of the Laplacian of e (the artificial
// transport's fourth-order drivers).
#include "st_pass.hpp"

namespace st {

// ---------------------------------------------------------------------
// D2 along the sweep of |S| and of Lap e, summed over the three directions
// into L2S = Lap|S| and L4e = Lap(Lap e) (src/hyper_dynamic.f:272-276).
// ACC = false overwrites (the first direction), true adds to the sums of
// the earlier directions (prefetched).  The mixed second derivatives of the
// viscous terms are done in the z and y first-derivative passes.
// ---------------------------------------------------------------------
template <bool ACC>
struct PassLapSE {
    static constexpr int NSLOT = 2, NOUT = 0, NEARLY = 2;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = ACC ? 2 : 0;
    DEV static void prefetch(const Geo &, const PassArgs &a, long gi, double *p)
    {
        if (ACC) {
            p[0] = a.L2S[gi];
            p[1] = a.L4e[gi];
        }
    }
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[2])
    {
        v[0] = a.S[gi];
        v[1] = a.L2e[gi];
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
    {
#pragma unroll 1
        for (int m = 0; m < 2; m++) {
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
        if constexpr (ACC) {
            a.L2S[gi] = v[2] + v[0];
            a.L4e[gi] = v[3] + v[1];
        } else {
            a.L2S[gi] = v[0];
            a.L4e[gi] = v[1];
        }
    }
};


hipError_t launch_lapse(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    if (kind != K_LAPSE) return hipErrorInvalidValue;
    if (a.first) return launch<PassLapSE<false>>(t, nt, dir, My, Mz, Mx, a, s);
    return launch<PassLapSE<true>>(t, nt, dir, My, Mz, Mx, a, s);
}

hipError_t launch_lapse_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    if (kind != K_LAPSE) return hipErrorInvalidValue;
    if (sp.a[sp.first].first) return launch_ragged<PassLapSE<false>>(sp, s);
    return launch_ragged<PassLapSE<true>>(sp, s);
}

}  // namespace st
