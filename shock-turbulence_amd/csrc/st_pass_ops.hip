// Single-operator passes: the DFDX/D2FDX2/Compact_Filter/Gaussian entry points.
#include "st_pass.hpp"

namespace st {

// ---------------------------------------------------------------------
// single-operator passes (the DFDX/D2FDX2/Compact_Filter_X/Gaussian_X
// entry points of the C-ABI)
// ---------------------------------------------------------------------
struct PassD1 : NoReuse {
    static constexpr int NSLOT = 1, NOUT = 1, NEARLY = 0;
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[1]) { v[0] = a.in0[gi]; }
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[1][M])
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
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[1][M])
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
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[1][M])
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
    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[1][M])
    {
        double f[M], w[M + 2 * HALO];
#pragma unroll
        for (int k = 0; k < M; k++) f[k] = L.ld(0, k);
        window<M>(f, L, w);
        op_gauss<M>(w, a.C, out[0]);
    }
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v) { a.out0[gi] = v[0]; }
};


hipError_t launch_ops(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_D1: return launch<PassD1>(t, nt, dir, My, Mz, Mx, a, s);
    case K_D2: return launch<PassD2>(t, nt, dir, My, Mz, Mx, a, s);
    case K_FILT: return launch<PassFilt>(t, nt, dir, My, Mz, Mx, a, s);
    case K_GAUSS: return launch<PassGauss>(t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st
