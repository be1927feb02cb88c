// The last pass of the artificial-transport divergence fused with the
// low-storage RK update and the first pass of the compact filter.
#include "st_pass_hdiv.hpp"

namespace st {

// ---------------------------------------------------------------------
// The last direction of the divergence (src/hyper_rhs.f:159-491) is the
// first of the filter (the stage runs the divergence z, x, y and the filter
// y, z, x), so one sweep ends the right-hand side and starts the stage
// update (src/lddrk4_periodic_sf.f:344-356):
//   compute   D1 of the staged fluxes (as the hyper_div pass)
//   mid       per point: dQ = dQ + D1 terms (continuity unchanged), the
//             sponge term when one is set, QW = a QW + dt dQ,
//             Q* = Q + b QW (src/lddrk4_periodic_sf.f:349-354), Q* staged
//             in the five slots
//   compute2  filter of Q* along the sweep (Compact_Filter_X/Y/Z,
//             src/filters_periodic.f:115,254,393)
//   store     Q = filtered Q*
// The right-hand side itself is not written back: nothing reads it after
// the stage update.  Rounding is that of the two passes run apart.
// ---------------------------------------------------------------------
struct PassHdivRkf {
    static constexpr int NSLOT = 5, NOUT = 0, NEARLY = 5;
    static constexpr bool REUSE = true;
    static constexpr bool MID = true;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    // the momenta and density that the fluxes read go on to the update in
    // registers (mid reads only the energy of Q)
    static constexpr int NLCARRY = 4;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &g, const PassArgs &a, long gi, double (&v)[5 + NLCARRY])
    {
        double t[4];
        PassHyperDiv::load(g, a, gi, t);
#pragma unroll
        for (int c = 0; c < 4; c++) v[c] = t[c];
        v[4] = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) v[5 + c] = a.Q[c * a.N + gi];
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &g, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&o)[1][M])
    {
        PassHyperDiv::compute<M, P, NP>(g, a, L, o);
    }
    DEV static void mid(const Geo &g, const PassArgs &a, long gi, int ix, double (&v)[5], const double *lq)
    {
        const long N = a.N;
        double dq[5];
        dq[0] = a.dQ[gi] + v[0];
        dq[1] = a.dQ[N + gi] + v[1];
        dq[2] = a.dQ[2 * N + gi] + v[2];
        dq[3] = a.dQ[3 * N + gi];
        dq[4] = a.dQ[4 * N + gi] + v[3];
        // Sponge_Forcing after Hyper_RHS (src/lddrk4_nonperiodic_mf.f:462-466),
        // as k_sponge rounds it; sigma and the reference state by x index
        if (a.spg && g.dir != 0) ix = (int)((gi % a.ni) / ((long)g.My * g.Mz));
        const double sg = a.spg ? a.spg[ix] : 0.0;
        const double *ref = a.spg ? a.spg + g.Mx + (ix < a.spg_left ? 0 : 5) : nullptr;
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const long o = c * N + gi;
            const double q = c < 4 ? lq[c] : a.Q[o];
            if (sg != 0.0) dq[c] = fma(-sg, q - ref[c], dq[c]);
            // as PassFilterQ<0> (alpha = 0: QW not read)
            const double qw = a.rk_a == 0.0 ? fma(a.dt, dq[c], 0.0) : fma(a.dt, dq[c], a.rk_a * a.QW[o]);
            a.QW[o] = qw;
            v[c] = fma(a.rk_b, qw, q);
        }
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute2(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L)
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
#pragma unroll
        for (int c = 0; c < 5; c++) a.Q[c * a.N + gi] = v[c];
    }
};


hipError_t launch_hdivrk(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a,
                         hipStream_t s)
{
    if (kind != K_HDIV_RKF || dir == 2) return hipErrorInvalidValue;
    return launch<PassHdivRkf>(t, nt, dir, My, Mz, Mx, a, s);
}

hipError_t launch_hdivrk_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    if (kind != K_HDIV_RKF || sp.dir == 2) return hipErrorInvalidValue;
    return launch_ragged<PassHdivRkf>(sp, s);
}

}  // namespace st
