// Stage-update and compact-filter passes.
#include "st_pass.hpp"

namespace st {

// ---------------------------------------------------------------------
// Low-storage RK update fused into the first pass of the compact filter
// (src/lddrk4_periodic_sf.f:349-354), then the other two filter passes; the
// last applies Fix_Pressure (src/lddrk4_periodic_sf.f:436-473).  The stage
// filters y, z, x: the three line filters commute, so the order changes
// only the rounding.
// ---------------------------------------------------------------------
template <int KIND>   // 0: RK update + filter, 1: filter, 2: filter + pressure fix
struct PassFilterQ {
    static constexpr int NSLOT = 5, NOUT = 0, NEARLY = 5;
    static constexpr bool REUSE = true;
    static constexpr int MIN_WAVES = 1;
    // the RK update loads 15 fields per point: two points in flight at once
    static constexpr int LOAD_UNROLL = KIND == 0 ? 1 : 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[5])
    {
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const long o = c * a.N + gi;
            if (KIND == 0) {
                // contractions written out, so that the fused form of this
                // update (PassHdivRkf) rounds alike; the first stage of a
                // loop (alpha = 0) does not read QW, so the loop needs no
                // clearing of it
                const double qw = a.rk_a == 0.0 ? fma(a.dt, a.dQ[o], 0.0) : fma(a.dt, a.dQ[o], a.rk_a * a.QW[o]);
                a.QW[o] = qw;
                v[c] = fma(a.rk_b, qw, a.Q[o]);
            } else {
                v[c] = a.Q[o];
            }
        }
    }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&)[1][M])
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
            a.Q[gi] = v[0];
            a.Q[N + gi] = v[1];
            a.Q[2 * N + gi] = v[2];
            a.Q[3 * N + gi] = v[3];
            a.Q[4 * N + gi] = fixed_energy(a.ph, v[0], v[1], v[2], v[3], v[4]);
            return;
        }
#pragma unroll
        for (int c = 0; c < 5; c++) a.Q[c * N + gi] = v[c];
    }
};


hipError_t launch_filter(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_RKF: return launch<PassFilterQ<0>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_FQ: return launch<PassFilterQ<1>>(t, nt, dir, My, Mz, Mx, a, s);
    case K_FQ_FIX: return launch<PassFilterQ<2>>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_filter_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    switch (kind) {
    case K_RKF: return launch_ragged<PassFilterQ<0>>(sp, s);
    case K_FQ: return launch_ragged<PassFilterQ<1>>(sp, s);
    case K_FQ_FIX: return launch_ragged<PassFilterQ<2>>(sp, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st
