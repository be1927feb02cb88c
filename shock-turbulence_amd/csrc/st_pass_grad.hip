// First-derivative passes of the right-hand side.
#include "st_pass.hpp"

namespace st {

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
// 2 for the y pass, whose nine staged fields allow two 256-thread
// workgroups per CU anyway, and for the z pass at 8 points per lane (two
// 8-line workgroups per CU); otherwise the compiler's own choice.  LEAN
// selects the operator order that keeps fewer values live (the x pass at
// every tile width, the z pass at 1024 threads and at 8 points per lane).
// CROSS (y lines, after the x and z passes have left their partial sums):
// the store phase also assembles the right-hand side of the first-derivative
// part -- the y terms, the z and x partials and the viscous terms built from
// products of first derivatives -- and |S| and C_beta, from the derivatives
// of all three directions at the point (the y ones carried from the early
// store in registers).
// NMX: stored first derivatives staged as well, whose D1 along the sweep
// gives the mixed second derivatives of the viscous terms
// (src/rhsv_periodic_sf.f:1062-1070,1281-1363):
//   y (4): D1y of du/dx, dv/dx, dv/dz, dw/dz;  z (2): D1z of du/dx, dw/dx
// each times (1/3 + bulk) mu into the momentum component it belongs to and,
// times that velocity, into the energy.  The x sweep runs first, so the z
// and y sweeps find the derivatives they need.
// TWO (the z sweep at 8 points per lane): the two mixed inputs are staged in
// the slots of v and rho, whose values wait in registers (NHOLD) until the
// mixed terms are done (compute_a); five slots instead of seven, so that
// 8-line tiles fit two workgroups per CU (0.70 -> 0.645 ms at 256^3; at 4
// points per lane the extra phase costs more than it gives, 0.70 -> 0.79).
template <int MW, bool LEAN, bool CROSS, int NMX, bool TWO_ = false>
struct PassGrad {
    static constexpr bool TWO = TWO_;
    static_assert(!TWO || NMX == 2, "two-round staging is the z sweep's");
    static constexpr int NSLOT = TWO ? 5 : 5 + NMX, NOUT = 5, NEARLY = 5;
    static constexpr int NHOLD = TWO ? 2 : 0;
    // CROSS: u, v, w and T of each point go from the load phase to the late
    // store in registers (the store assembles the right-hand side from them)
    static constexpr int NLCARRY = CROSS ? 4 : 0;
    HD static constexpr int hold_slot(int j) { return j == 0 ? 1 : 4; }
    static constexpr bool REUSE = true;
    static constexpr int NCARRY = CROSS ? 4 : 0;
    // CROSS: the staged inputs of the mixed terms (du/dx, dv/dx, dv/dz,
    // dw/dz) reach the store after the late results
    static constexpr int NKEEP = CROSS ? NMX : 0;
    static constexpr int MIN_WAVES = MW;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    // the additive LDS layout for the x and y sweeps (fewer address
    // registers: the x pass at 8 points per lane no longer spills, and the y
    // pass runs 2% faster) and for z at 8 points per lane; the swizzled one
    // for z at 4 points per lane, whose 4-line tiles would take 4-way bank
    // conflicts in the staging
    HD static constexpr bool lds_add(int m) { return NMX != 2 || m >= 8; }
    // CROSS: the store phases' global reads (sfetch_early: the z and x
    // partials of the Laplacian of e; sfetch: four x and z derivatives, and
    // the z and x partial sums) go out for all the thread's points (two at a
    // time) before the first store
    static constexpr int NSFE = CROSS ? 2 : 0, NSF = CROSS ? 14 : 0, SF_BATCH = 2;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    // G index of mixed input m: y: gxx=0, gyx=1, gyz=7, gzz=8; z: gxx=0, gzx=2
    HD static constexpr int msrc(int m) { return NMX == 4 ? (m == 0 ? 0 : m == 1 ? 1 : m == 2 ? 7 : 8) : (m == 0 ? 0 : 2); }
    // momentum component that D1 of mixed input m feeds
    HD static constexpr int mdst(int m) { return NMX == 4 ? (m == 1 ? 0 : m == 2 ? 2 : 1) : (m == 0 ? 2 : 0); }
    // slot of mixed input m
    HD static constexpr int mslot(int m) { return TWO ? (m == 0 ? 1 : 4) : 5 + m; }
    DEV static void load(const Geo &, const PassArgs &a, long gi, double (&v)[NSLOT + NHOLD + NLCARRY])
    {
        const Prim pr = prim(a.ph, a.Q[gi], a.Q[a.N + gi], a.Q[2 * a.N + gi], a.Q[3 * a.N + gi], a.Q[4 * a.N + gi]);
        v[0] = pr.u;
        v[2] = pr.w;
        v[3] = pr.T;
        if constexpr (CROSS) {
            v[NSLOT] = pr.u;
            v[NSLOT + 1] = pr.v;
            v[NSLOT + 2] = pr.w;
            v[NSLOT + 3] = pr.T;
        }
        if constexpr (TWO) {
            v[5] = pr.v;
            v[6] = a.Q[3 * a.N + gi];
        } else {
            v[1] = pr.v;
            v[4] = a.Q[3 * a.N + gi];
        }
#pragma unroll
        for (int m = 0; m < NMX; m++) v[mslot(m)] = a.G[msrc(m) * a.N + gi];
    }
    // The mixed-derivative terms, first (the velocities are still staged).
    template <int M, int P, bool NP, bool AL>
    DEV static void compute_mix(const PassArgs &a, const Lane<M, P, NP, AL> &L, const double (&mu)[M], double (&out)[5][M])
    {
        const double cb = a.ph.bulk + 1.0 / 3.0;
#pragma unroll 1
        for (int m = 0; m < NMX; m++) {
            const int comp = mdst(m);
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(mslot(m), k);
            window<M>(f, L, w);
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) {
                const double t = (cb * mu[k]) * r[k];
                acc<0x07, M>(out, comp, k, t);
                out[4][k] += L.ld(comp, k) * t;
            }
        }
    }
    // TWO: the mixed terms, while their inputs hold the slots of v and rho.
    template <int M, int P, bool NP, bool AL>
    DEV static void compute_a(const Geo &, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[5][M])
    {
        double mu[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            mu[k] = visc(a.ph, L.ld(3, k));
#pragma unroll
            for (int i = 0; i < 5; i++) out[i][k] = 0.0;
        }
        compute_mix<M, P, NP>(a, L, mu, out);
    }
    // Accumulate into row i of out (one of the rows in MASK) with static
    // register indices: every row in MASK takes w_r v, w_r = 1 for row i and
    // 0 otherwise (i is the same on every lane, so w_r is a scalar and each
    // row costs one fused multiply-add, where a select costs three).
    template <int MASK, int M>
    DEV static void acc(double (&out)[5][M], int i, int k, double v)
    {
#pragma unroll
        for (int r = 0; r < 5; r++)
            if ((MASK >> r) & 1) out[r][k] = fma(i == r ? 1.0 : 0.0, v, out[r][k]);
    }
    // Operator order: fluxes and the mass flux first, while u, v, w, T and
    // rho are all needed; then T, whose slot takes dT/dx_d and rho's slot
    // the Laplacian of e; then each velocity, whose slot takes its
    // derivative.  The loop bodies are free of branches so that the
    // accumulators stay in place across iterations.
    template <int M, int P, bool NP, bool AL>
    DEV static void compute_wide(const Geo &g, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[5][M])
    {
        const Phys &ph = a.ph;
        const int d = g.dir;
        const double pfac = ph.gm1 * ph.ginv;
        double mu[M], rud[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            mu[k] = visc(ph, L.ld(3, k));
            rud[k] = -0.5 * L.ld(4, k) * L.ld(d, k);
            if constexpr (!TWO) {
#pragma unroll
                for (int i = 0; i < 5; i++) out[i][k] = 0.0;
            }
        }
        if constexpr (NMX > 0 && !TWO) compute_mix<M, P, NP>(a, L, mu, out);
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

    // Operator order: fluxes and the mass flux first, while u, v, w, T and
    // rho are all needed; then the velocities (u_d last), whose slots take
    // their derivatives; then T, whose slot takes dT/dx_d and rho's slot
    // the Laplacian of e.  The loop bodies are free of branches so that the
    // accumulators stay in place across iterations, and each loop iteration
    // keeps one window and one solve live at a time.
    template <int M, int P, bool NP, bool AL>
    DEV static void compute_lean(const Geo &g, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[5][M])
    {
        const Phys &ph = a.ph;
        const int d = g.dir;
        const double pfac = ph.gm1 * ph.ginv;
        double mu[M];
        if constexpr (!TWO) {
#pragma unroll
            for (int k = 0; k < M; k++) {
#pragma unroll
                for (int i = 0; i < 5; i++) out[i][k] = 0.0;
            }
            if constexpr (NMX > 0) {
#pragma unroll
                for (int k = 0; k < M; k++) mu[k] = visc(ph, L.ld(3, k));
                compute_mix<M, P, NP>(a, L, mu, out);
            }
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
        // the viscosity, first needed here (without mixed terms, or with
        // them done in compute_a) so that it is not held through the flux
        // solves
        if constexpr (NMX == 0 || TWO) {
#pragma unroll
            for (int k = 0; k < M; k++) mu[k] = visc(ph, L.ld(3, k));
        }
        // u, v, w (u_d last): second, then first derivatives
#pragma unroll 1
        for (int it = 0; it < 3; it++) {
            const int comp = it == 2 ? d : (d + 1 + it) % 3;
            const double wd = comp == d ? (4.0 / 3.0 + ph.bulk) : 1.0;
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(comp, k);
            window<M>(f, L, w);
            double vt[M];
            op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) {
                vt[k] = wd * mu[k] * r[k];
                out[4][k] += w[HALO + k] * vt[k];
            }
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) {
                const double rud = -0.5 * L.ld(4, k) * L.ld(d, k);
                acc<0x07, M>(out, comp, k, vt[k] + rud * r[k]);
                L.st(comp, k, r[k]);
            }
        }
        // T: first and second derivatives; the Laplacian of e = T/gamma is
        // D2(T)/gamma (the operator is linear)
        {
            double f[M], w[M + 2 * HALO], r[M];
#pragma unroll
            for (int k = 0; k < M; k++) f[k] = L.ld(3, k);
            window<M>(f, L, w);
            op_d2<M>(w, a.C, a.K2, a.tab->op[1], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) {
                L.st(4, k, r[k] * ph.ginv);
                out[4][k] += ph.pr_inv * mu[k] * r[k];
            }
            op_d1<M>(w, a.C, a.K1, a.tab->op[0], L, r);
#pragma unroll
            for (int k = 0; k < M; k++) {
                L.st(3, k, r[k]);
                const double mup = dvisc(ph, w[HALO + k], mu[k]);
                out[4][k] += ph.pr_inv * (mup * r[k] * r[k]);
            }
        }
    }

    template <int M, int P, bool NP, bool AL>
    DEV static void compute(const Geo &g, const PassArgs &a, const Lane<M, P, NP, AL> &L, double (&out)[5][M])
    {
        if constexpr (LEAN) compute_lean<M, P, NP>(g, a, L, out);
        else compute_wide<M, P, NP>(g, a, L, out);
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

    // CROSS: G(., y) and dT/dy go out and stay in the carry; the Laplacian
    // of e adds the z, then the x partial.
    DEV static void sfetch_early(const Geo &, const PassArgs &a, long gi, double *sf)
    {
        sf[0] = a.L2e ? a.L2ep[gi] : 0.0;
        sf[1] = a.L2e ? a.L2ep[a.N + gi] : 0.0;
    }
    DEV static void store_early(const Geo &, const PassArgs &a, long gi, const double *v, double *cr,
                                const double *sf)
    {
        const long N = a.N;
#pragma unroll
        for (int c = 0; c < 3; c++) a.G[(3 + c) * N + gi] = v[c];
        a.G[10 * N + gi] = v[3];
#pragma unroll
        for (int c = 0; c < 4; c++) cr[c] = v[c];
        if (a.L2e) a.L2e[gi] = (v[4] + sf[0]) + sf[1];
    }
    // the store phase's reads: du/dz, dw/dx, dT/dx, dT/dz, the z partials
    // (5), the x partials (5)
    DEV static void sfetch(const Geo &, const PassArgs &a, long gi, double *sf)
    {
        const long N = a.N;
        sf[0] = a.G[2 * N + gi];
        sf[1] = a.G[6 * N + gi];
        sf[2] = a.G[9 * N + gi];
        sf[3] = a.G[11 * N + gi];
#pragma unroll
        for (int c = 0; c < 10; c++) sf[4 + c] = a.dQp[c * N + gi];
    }
    // CROSS: viscous terms built from products of first derivatives
    // (src/rhsv_periodic_sf.f:1128-1154 without the mixed and own-direction
    // second derivatives, and :1252-1275), plus |S| (src/hyper_dynamic.f:
    // 12-33, 227-245) and the dilatation sensor C_beta
    // (src/hyper_dynamic.f:142-160; skipped when Cb is null: the
    // constant-coefficient src/hyper_routines.f); |S| and C_beta only with
    // the artificial transport on (L2e set).
    DEV static void store(const Geo &, const PassArgs &a, long gi, const double *v, const double *cr,
                          const double *sf, const double *lc)
    {
        const long N = a.N;
        const Phys &ph = a.ph;
        struct { double u, v, w, T; } pr = {lc[0], lc[1], lc[2], lc[3]};
        static_assert(!CROSS || (NMX == 4 && msrc(0) == 0 && msrc(1) == 1 && msrc(2) == 7 && msrc(3) == 8),
                      "the kept slots are du/dx, dv/dx, dv/dz, dw/dz");
        const double gxx = v[5], gyx = v[6], gzx = sf[0];
        const double gxy = cr[0], gyy = cr[1], gzy = cr[2];
        const double gxz = sf[1], gyz = v[7], gzz = v[8];
        const double mu = visc(ph, pr.T), mup = dvisc(ph, pr.T, mu);
        const double tx = mup * sf[2], ty = mup * cr[3], tz = mup * sf[3];
        const double f1 = ph.bulk - 2.0 / 3.0;
        const double dil = gxx + gyy + gzz;
        const double t1 = gyx + gxy, t2 = gzx + gxz, t3 = gzy + gyz;
        const double fX = tx * (2.0 * gxx + f1 * dil) + ty * t1 + tz * t2;
        const double fY = ty * (2.0 * gyy + f1 * dil) + tx * t1 + tz * t3;
        const double fZ = tz * (2.0 * gzz + f1 * dil) + tx * t2 + ty * t3;
        const double *Pz = sf + 4, *Px = sf + 9;
        a.dQ[gi] = ((v[0] + Pz[0]) + Px[0]) + fX;
        a.dQ[N + gi] = ((v[1] + Pz[1]) + Px[1]) + fY;
        a.dQ[2 * N + gi] = ((v[2] + Pz[2]) + Px[2]) + fZ;
        a.dQ[3 * N + gi] = (v[3] + Pz[3]) + Px[3];
        a.dQ[4 * N + gi] = ((v[4] + Pz[4]) + Px[4])
                           + (pr.u * fX + pr.v * fY + pr.w * fZ
                              + mu * (2.0 * (gxx * gxx + gyy * gyy + gzz * gzz) + f1 * dil * dil)
                              + mu * (t1 * t1 + t2 * t2 + t3 * t3));
        if (a.L2e) {
            const double s12 = 0.5 * (gxy + gyx), s13 = 0.5 * (gxz + gzx), s23 = 0.5 * (gyz + gzy);
            a.S[gi] = sqrt(0.5 * (gxx * gxx + gyy * gyy + gzz * gzz + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
            if (a.Cb) {
                const double wx = gzy - gyz, wy = gxz - gzx, wz = gyx - gxy;
                const double om2 = wx * wx + wy * wy + wz * wz;
                const double fac = 0.5 * (1.0 - tanh(2.5 + 10.0 * dil));
                a.Cb[gi] = fac * dil * dil / (dil * dil + om2 + 1.0e-7) + 1.0e-20;
            }
        }
    }
};


hipError_t launch_grad(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    // x first (no mixed terms), z with two, y last with four and the
    // pointwise assembly
    if (kind == K_GRAD && dir == 0) {
        return launch<PassGrad<1, true, false, 0>>(t, nt, dir, My, Mz, Mx, a, s);
    }
    if (kind == K_GRAD && dir == 2) {
        // 8 points per lane on 256 threads: the lean order within two waves per SIMD
        if (nt == 256 && t.M == 8 && t.P == 32) return launch_at<PassGrad<2, true, false, 2, true>, 8, 32, 256>(t, dir, My, Mz, Mx, a, s);
        if (nt == 1024) return launch<PassGrad<1, true, false, 2>, false>(t, nt, dir, My, Mz, Mx, a, s);
        return launch<PassGrad<1, false, false, 2>, false>(t, nt, dir, My, Mz, Mx, a, s);
    }
    if (kind == K_GRAD_CROSS && dir == 1) {
        if (nt == 1024) return launch<PassGrad<1, true, true, 4>, false>(t, nt, dir, My, Mz, Mx, a, s);
        return launch<PassGrad<2, false, true, 4>, false>(t, nt, dir, My, Mz, Mx, a, s);
    }
    return hipErrorInvalidValue;
}

// The 256-thread variants (those of the plain path below 1024 threads).
hipError_t launch_grad_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    if (kind == K_GRAD && sp.dir == 0) return launch_ragged<PassGrad<1, true, false, 0>>(sp, s);
    if (kind == K_GRAD && sp.dir == 2) return launch_ragged<PassGrad<1, false, false, 2>>(sp, s);
    if (kind == K_GRAD_CROSS && sp.dir == 1) return launch_ragged<PassGrad<2, false, true, 4>>(sp, s);
    return hipErrorInvalidValue;
}

}  // namespace st
