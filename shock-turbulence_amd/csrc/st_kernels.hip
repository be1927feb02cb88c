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

void pass_default_cfg(int kind, int dir, int *narrow, int *nt)
{
    // Chosen by st_autotune at 256^3 (per pass and direction): wider tiles
    // (more lines, so longer unit-stride rows in the strided x and z sweeps)
    // where the pass's registers allow the waves; M/2 points per lane
    // ("narrow") where the pass holds many fields in registers.
    struct C { int n, t; };
    C c = {0, 256};
    switch (kind) {
    case K_GRAD: c = {0, 256}; break;
    case K_GRAD_CROSS: c = {1, 256}; break;
    case K_LAPSE: c = dir == 1 ? C{1, 512} : C{0, 512}; break;
    case K_LAP2: c = {0, 512}; break;
    case K_LAP2_LAST: c = {1, 256}; break;
    case K_LAPSE_LAP2: c = {0, 512}; break;
    case K_GAUSS_ETA: c = dir == 0 ? C{1, 1024} : C{1, 256}; break;
    case K_GAUSS_ETA_LAST: c = {1, 1024}; break;
    case K_HDIV: c = dir == 0 ? C{0, 512} : C{1, 512}; break;
    case K_HDIV_RKF: c = dir == 0 ? C{0, 512} : C{1, 256}; break;
    case K_RKF: c = {0, 256}; break;
    case K_FQ: c = {0, 256}; break;
    case K_FQ_FIX: c = {0, 256}; break;
    default: break;
    }
    *narrow = c.n;
    *nt = c.t;
}

hipError_t run_pass(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    switch (kind) {
    case K_D1: case K_D2: case K_FILT: case K_GAUSS:
        return launch_ops(kind, t, nt, dir, My, Mz, Mx, a, s);
    case K_GRAD: case K_GRAD_CROSS:
        return launch_grad(kind, t, nt, dir, My, Mz, Mx, a, s);
    case K_LAPSE:
        return launch_lapse(kind, t, nt, dir, My, Mz, Mx, a, s);
    case K_LAP2: case K_LAP2_LAST: case K_GAUSS_ETA: case K_GAUSS_ETA_LAST: case K_HDIV: case K_LAPSE_LAP2:
        return launch_hyper(kind, t, nt, dir, My, Mz, Mx, a, s);
    case K_RKF: case K_FQ: case K_FQ_FIX:
        return launch_filter(kind, t, nt, dir, My, Mz, Mx, a, s);
    case K_HDIV_RKF:
        return launch_hdivrk(kind, t, nt, dir, My, Mz, Mx, a, s);
    default: return hipErrorInvalidValue;
    }
}

int ragged_member(int n)
{
    for (int c = 0; c < 4; c++)
        if (RAGGED_N[c] == n) return c;
    return -1;
}

hipError_t run_pass_ragged(int kind, const RaggedSpec &sp, hipStream_t s)
{
    switch (kind) {
    case K_GRAD: case K_GRAD_CROSS:
        return launch_grad_ragged(kind, sp, s);
    case K_LAPSE:
        return launch_lapse_ragged(kind, sp, s);
    case K_LAP2: case K_LAP2_LAST: case K_GAUSS_ETA: case K_GAUSS_ETA_LAST: case K_HDIV: case K_LAPSE_LAP2:
        return launch_hyper_ragged(kind, sp, s);
    case K_RKF: case K_FQ: case K_FQ_FIX:
        return launch_filter_ragged(kind, sp, s);
    case K_HDIV_RKF:
        return launch_hdivrk_ragged(kind, sp, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace st

// ---------------------------------------------------------------------
// reductions: one partial per block and slot, partial[(slot*nb + block)*NV + v]
// (slot: the slot number, not its index in the launch)
// ---------------------------------------------------------------------
namespace st {

template <int NV, class OP>
DEV void block_reduce(double (&v)[NV], OP op, double *partial, int slot)
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
        partial[((long)slot * gridDim.x + blockIdx.x) * NV + threadIdx.x] = x;
    }
}

struct OpMax { DEV double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpMin { DEV double operator()(double a, double b) const { return a < b ? a : b; } };
struct OpSum { DEV double operator()(double a, double b) const { return a + b; } };

// Compute_CFL src/lddrk4_periodic_sf.f:29-56
__global__ void __launch_bounds__(256) k_cfl(PassArgs a, double *partial)
{
    const int slot = a.slots ? a.slots[blockIdx.y] : blockIdx.y;
    if (a.act && !a.act[slot]) return;
    const long N = a.N, off = (long)slot * a.ni;
    double v[2] = {-INFINITY, -INFINITY};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.ni; i += (long)gridDim.x * blockDim.x) {
        const long gi = off + i;
        double w1, w2;
        cfl_point(a.ph, a.Q[gi], a.Q[N + gi], a.Q[2 * N + gi], a.Q[3 * N + gi], a.Q[4 * N + gi], w1, w2);
        v[0] = w1 > v[0] ? w1 : v[0];
        v[1] = w2 > v[1] ? w2 : v[1];
    }
    block_reduce<2>(v, OpMax(), partial, slot);
}

// Compute_Tstep src/lddrk4_periodic_sf.f:105-132 (beta*, mu*, kappa* in Hc)
__global__ void __launch_bounds__(256) k_tstep(PassArgs a, double *partial)
{
    const int slot = a.slots ? a.slots[blockIdx.y] : blockIdx.y;
    if (a.act && !a.act[slot]) return;
    const long N = a.N, off = (long)slot * a.ni;
    const Phys &ph = a.ph;
    const double fac = ph.gamma * ph.gm1;
    const double del = ph.D[0];
    double v[3] = {INFINITY, INFINITY, INFINITY};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.ni; i += (long)gridDim.x * blockDim.x) {
        const long gi = off + i;
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
    block_reduce<3>(v, OpMin(), partial, slot);
}

// Compute_Mom/KE/Mass/Energy/Entropy src/lddrk4_periodic_sf.f:153-302 (sums)
__global__ void __launch_bounds__(256) k_diag(PassArgs a, double *partial)
{
    const int slot = a.slots ? a.slots[blockIdx.y] : blockIdx.y;
    if (a.act && !a.act[slot]) return;
    const long N = a.N, off = (long)slot * a.ni;
    const Phys &ph = a.ph;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.ni; i += (long)gridDim.x * blockDim.x) {
        const long gi = off + i;
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
    block_reduce<7>(v, OpSum(), partial, slot);
}

__global__ void __launch_bounds__(256) k_fill(double *p, long n, double v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// p[f*N + s*ni + i] = v for the launch's slots s (blockIdx.y = f*nslot + index)
__global__ void __launch_bounds__(256) k_fill_slots(double *p, long N, long ni, int nslot, const int *slots,
                                                     const int *act, double v)
{
    const int f = blockIdx.y / nslot, si = blockIdx.y - f * nslot;
    const int slot = slots ? slots[si] : si;
    if (act && !act[slot]) return;
    double *q = p + f * N + (long)slot * ni;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) q[i] = v;
}

// Fix_Pressure on a resident array (src/lddrk4_periodic_sf.f:436-473), slot 0
__global__ void __launch_bounds__(256) k_fixp(PassArgs a, double *Qx)
{
    const long N = a.N;
    const Phys &ph = a.ph;
    for (long gi = blockIdx.x * (long)blockDim.x + threadIdx.x; gi < a.ni; gi += (long)gridDim.x * blockDim.x) {
        Qx[4 * N + gi] = fixed_energy(ph, Qx[gi], Qx[N + gi], Qx[2 * N + gi], Qx[3 * N + gi], Qx[4 * N + gi]);
    }
}

// Device-side Compute_Tstep for every slot (one thread per slot): fold the
// CFL and dt-limit partials, dt = cfl_cal/cfl_resv, then the hyper limits
// when the slot's step passes the driver's rule (src/lddrk4_periodic_sf.f:96-142;
// first_step = 1 applies them from the first step,
// src/lddrk4_nonperiodic_sf.f:105).  stepc is the driver's step counter
// (1 before the first step).  A slot runs this step (run = 1) while it is
// active and its counter is below its target.
__global__ void k_dt(const double *pc, const double *pt, int nb, int nslot, const int *active, const int *target,
                     const int *stepc, double cfl_cal, int hyper, int first_step, int *run, double *dt, double *dt_last)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslot) return;
    const int go = active[slot] && stepc[slot] < target[slot];
    run[slot] = go;
    if (!go) return;
    double resv = -INFINITY;
    for (int b = 0; b < nb; b++) {
        const double x = pc[((long)slot * nb + b) * 2 + 1];
        resv = x > resv ? x : resv;
    }
    double t = cfl_cal / resv;
    if (hyper == 1 && (stepc[slot] > 1 || first_step)) {
        double v[3] = {INFINITY, INFINITY, INFINITY};
        for (int b = 0; b < nb; b++)
            for (int i = 0; i < 3; i++) {
                const double x = pt[((long)slot * nb + b) * 3 + i];
                v[i] = x < v[i] ? x : v[i];
            }
        const double p2 = (double)0.2f;                 // default-real literal 0.2
        if (p2 * v[0] < t) t = p2 * v[0];
        if (p2 * v[1] < t) t = p2 * v[1];
        if (p2 * v[2] < t) t = p2 * v[2];
    }
    dt[slot] = t;
    dt_last[slot] = t;
}

// time += dt after each RK loop; step += 1 after the second
// (src/lddrk4_periodic_sf.f:372,419-420)
__global__ void k_clock(int nslot, const int *run, const double *dt, double *time, int *stepc, int bump)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslot || !run[slot]) return;
    time[slot] = time[slot] + dt[slot];
    if (bump) stepc[slot] = stepc[slot] + 1;
}

// Sponge_Forcing (src/sponge1d.f:85-95): dQ -= sigma(x) (Q - Qref) with
// sigma per x index in spg[0..Mx) and the reference states of the left
// and right layers in spg[Mx..Mx+5) and spg[Mx+5..Mx+10).
__global__ void __launch_bounds__(256) k_sponge(PassArgs a, const double *spg, int Mx, int nleft)
{
    const int slot = a.slots ? a.slots[blockIdx.y] : blockIdx.y;
    if (a.act && !a.act[slot]) return;
    const long N = a.N, off = (long)slot * a.ni, plane = a.ni / Mx;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.ni; i += (long)gridDim.x * blockDim.x) {
        const int ix = (int)(i / plane);
        const double sg = spg[ix];
        if (sg == 0.0) continue;
        const double *ref = spg + Mx + (ix < nleft ? 0 : 5);
        const long gi = off + i;
#pragma unroll
        for (int l = 0; l < 5; l++) a.dQ[l * N + gi] = fma(-sg, a.Q[l * N + gi] - ref[l], a.dQ[l * N + gi]);
    }
}

hipError_t run_sponge(const PassArgs &a, const double *spg, int Mx, int nleft, hipStream_t s)
{
    long blocks = (a.ni + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_sponge, dim3((unsigned)blocks, (unsigned)a.nslot), dim3(256), 0, s, a, spg, Mx, nleft);
    return hipGetLastError();
}

// ---------------------------------------------------------------------
// Compute_hyper outside a stage (src/hyper_dynamic.f:173-318; the static
// variant src/hyper_routines.f:66-214) and Compute_Grad (src/grad.f:13-115):
// the pointwise steps between single-operator passes.  Each acts on one
// slot: the pointers already carry the slot's offset, fields are N apart.
// ---------------------------------------------------------------------

// W[0..2] = Q_i * (1/rho) (src/hyper_dynamic.f:87-97, src/grad.f:40-48);
// with want_hyp also W[7] = |S| of the symmetrised /GRAD/ (0 when G is
// null: the coefficients of a start without gradients, src/main.f:100)
// (src/hyper_dynamic.f:227-245, :12-33) and W[8] = e (:251-266).
__global__ void __launch_bounds__(256) k_hyp_prep(const double *Q, const double *G, long N, long ni, int want_hyp,
                                                  double *W)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        const double q1 = Q[i], q2 = Q[N + i], q3 = Q[2 * N + i], q4 = Q[3 * N + i];
        const double t = 1.0 / q4;
        W[i] = q1 * t;
        W[N + i] = q2 * t;
        W[2 * N + i] = q3 * t;
        if (!want_hyp) continue;
        double s = 0.0;
        if (G) {
            const double s11 = G[i], s22 = G[4 * N + i], s33 = G[8 * N + i];
            const double s12 = 0.5 * (G[3 * N + i] + G[N + i]);
            const double s13 = 0.5 * (G[6 * N + i] + G[2 * N + i]);
            const double s23 = 0.5 * (G[7 * N + i] + G[5 * N + i]);
            s = sqrt(0.5 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        }
        W[7 * N + i] = s;
        W[8 * N + i] = t * (Q[4 * N + i] - 0.5 * t * (q1 * q1 + q2 * q2 + q3 * q3));
    }
}

// C_beta from the vorticity W[3..5] and dilatation W[6]
// (src/hyper_dynamic.f:142-160)
__global__ void __launch_bounds__(256) k_hyp_cb(const double *W, long N, long ni, double *Cb)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        const double g1 = W[3 * N + i], g2 = W[4 * N + i], g3 = W[5 * N + i], g4 = W[6 * N + i];
        const double fac = 0.5 * (1.0 - tanh(2.5 + 10.0 * g4));
        const double om2 = g1 * g1 + g2 * g2 + g3 * g3;
        Cb[i] = fac * g4 * g4 / (g4 * g4 + om2 + 1.0e-7) + 1.0e-20;
    }
}

// Compute_Laplacian's sum d2S = D2z + D2x + D2y (src/hyper_dynamic.f:54),
// optionally |.| (:278-279)
__global__ void __launch_bounds__(256) k_sum3(const double *z, const double *x, const double *y, long ni, int absval,
                                              double *out)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        const double s = z[i] + x[i] + y[i];
        out[i] = absval ? fabs(s) : s;
    }
}

// eta_1 = rho |Lap^2 S|, eta_2 = (rho/T) c |Lap^2 e| (src/hyper_dynamic.f:284-297)
__global__ void __launch_bounds__(256) k_hyp_eta(const double *Q, long N, long ni, Phys ph, const double *d4S,
                                                 const double *d4e, double *eta)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        const double q4 = Q[3 * N + i];
        const Prim pr = prim(ph, Q[i], Q[N + i], Q[2 * N + i], q4, Q[4 * N + i]);
        const double fac3 = (q4 / pr.T) * sqrt(ph.gamma * pr.p * pr.vs);
        eta[i] = q4 * d4S[i];
        eta[N + i] = fac3 * d4e[i];
    }
}

// beta*, mu*, kappa* from the smoothed eta (src/hyper_dynamic.f:303-314;
// C_beta = 1 without the sensor, src/hyper_routines.f:204)
__global__ void __launch_bounds__(256) k_hyp_coef(const double *g1, const double *g2, const double *Cb, long N,
                                                  long ni, Phys ph, double *Hc)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        Hc[i] = (Cb ? Cb[i] : 1.0) * g1[i] * ph.dx6;
        Hc[N + i] = ph.cmu * g1[i] * ph.dx6;
        Hc[2 * N + i] = ph.ck * g2[i] * ph.dx5;
    }
}

// Compute_Grad's reductions (src/grad.f:100-115): max |dil| and the sum of
// |omega|^2 over one slot, from /GRAD/ (G fields 0..8), partials per block.
__global__ void __launch_bounds__(256) k_gradred(const double *G, long N, long ni, double *partial)
{
    __shared__ double sh[2][4];
    double mx = 0.0, s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ni; i += (long)gridDim.x * blockDim.x) {
        const double dux_dx = G[i], duy_dx = G[N + i], duz_dx = G[2 * N + i];
        const double dux_dy = G[3 * N + i], duy_dy = G[4 * N + i], duz_dy = G[5 * N + i];
        const double dux_dz = G[6 * N + i], duy_dz = G[7 * N + i], duz_dz = G[8 * N + i];
        const double w1 = duz_dy - duy_dz, w2 = dux_dz - duz_dx, w3 = duy_dx - dux_dy;
        const double dil = fabs(dux_dx + duy_dy + duz_dz);
        mx = dil > mx ? dil : mx;
        s += w1 * w1 + w2 * w2 + w3 * w3;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_xor(mx, o);
        mx = a > mx ? a : mx;
        s += __shfl_xor(s, o);
    }
    if (lane == 0) {
        sh[0][wave] = mx;
        sh[1][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0][0], t = sh[1][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            m = sh[0][w] > m ? sh[0][w] : m;
            t += sh[1][w];
        }
        partial[2 * blockIdx.x] = m;
        partial[2 * blockIdx.x + 1] = t;
    }
}

static long pw_blocks(long ni)
{
    long b = (ni + 255) / 256;
    return b > 4096 ? 4096 : b;
}

hipError_t run_hyp_prep(const double *Q, const double *G, long N, long ni, int want_hyp, double *W, hipStream_t s)
{
    hipLaunchKernelGGL(k_hyp_prep, dim3((unsigned)pw_blocks(ni)), dim3(256), 0, s, Q, G, N, ni, want_hyp, W);
    return hipGetLastError();
}

hipError_t run_hyp_cb(const double *W, long N, long ni, double *Cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_hyp_cb, dim3((unsigned)pw_blocks(ni)), dim3(256), 0, s, W, N, ni, Cb);
    return hipGetLastError();
}

hipError_t run_sum3(const double *z, const double *x, const double *y, long ni, int absval, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum3, dim3((unsigned)pw_blocks(ni)), dim3(256), 0, s, z, x, y, ni, absval, out);
    return hipGetLastError();
}

hipError_t run_hyp_eta(const double *Q, long N, long ni, const Phys &ph, const double *d4S, const double *d4e,
                       double *eta, hipStream_t s)
{
    hipLaunchKernelGGL(k_hyp_eta, dim3((unsigned)pw_blocks(ni)), dim3(256), 0, s, Q, N, ni, ph, d4S, d4e, eta);
    return hipGetLastError();
}

hipError_t run_hyp_coef(const double *g1, const double *g2, const double *Cb, long N, long ni, const Phys &ph,
                        double *Hc, hipStream_t s)
{
    hipLaunchKernelGGL(k_hyp_coef, dim3((unsigned)pw_blocks(ni)), dim3(256), 0, s, g1, g2, Cb, N, ni, ph, Hc);
    return hipGetLastError();
}

int gradred_blocks(long ni)
{
    long b = (ni + 255) / 256;
    return (int)(b > 1024 ? 1024 : b);
}

hipError_t run_gradred(const double *G, long N, long ni, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(k_gradred, dim3((unsigned)gradred_blocks(ni)), dim3(256), 0, s, G, N, ni, partial);
    return hipGetLastError();
}

static long red_blocks(long N)
{
    long b = (N + 255) / 256;
    return b > 1024 ? 1024 : b;
}

hipError_t run_reduce(int which, const PassArgs &a, double *partial, int *nblocks, hipStream_t s)
{
    const long nb = red_blocks(a.ni);
    *nblocks = (int)nb;
    const dim3 grid((unsigned)nb, (unsigned)a.nslot);
    if (which == 0) hipLaunchKernelGGL(k_cfl, grid, dim3(256), 0, s, a, partial);
    else if (which == 1) hipLaunchKernelGGL(k_tstep, grid, dim3(256), 0, s, a, partial);
    else hipLaunchKernelGGL(k_diag, grid, dim3(256), 0, s, a, partial);
    return hipGetLastError();
}

int reduce_blocks(long ni) { return (int)red_blocks(ni); }

hipError_t run_fill(double *p, long n, double v, hipStream_t s)
{
    long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t run_fill_slots(double *p, int nfields, long N, long ni, int nslot, const int *slots, const int *act,
                          double v, hipStream_t s)
{
    long nb = (ni + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_fill_slots, dim3((unsigned)nb, (unsigned)(nfields * nslot)), dim3(256), 0, s, p, N, ni, nslot,
                       slots, act, v);
    return hipGetLastError();
}

hipError_t run_fixp(const PassArgs &a, double *Qx, hipStream_t s)
{
    const long nb = red_blocks(a.ni) * 4;
    hipLaunchKernelGGL(k_fixp, dim3((unsigned)nb), dim3(256), 0, s, a, Qx);
    return hipGetLastError();
}

hipError_t run_dt(const double *pc, const double *pt, int nb, int nslot, const int *active, const int *target,
                  const int *stepc, double cfl_cal, int hyper, int first_step, int *run, double *dt, double *dt_last,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_dt, dim3((unsigned)((nslot + 63) / 64)), dim3(64), 0, s, pc, pt, nb, nslot, active, target,
                       stepc, cfl_cal, hyper, first_step, run, dt, dt_last);
    return hipGetLastError();
}

hipError_t run_clock(int nslot, const int *run, const double *dt, double *time, int *stepc, int bump, hipStream_t s)
{
    hipLaunchKernelGGL(k_clock, dim3((unsigned)((nslot + 63) / 64)), dim3(64), 0, s, nslot, run, dt, time, stepc, bump);
    return hipGetLastError();
}

}  // namespace st

