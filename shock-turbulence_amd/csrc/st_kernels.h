// Kernel-launch interface between the C-ABI host code and the HIP passes.
#pragma once

#include <hip/hip_runtime.h>

#include "st_internal.h"

namespace st {

// Everything a pass may read or write.  Unused pointers stay null.
struct PassArgs {
    const DirTab *tab;    // device copy of this direction's tables
    Coef C;
    Phys ph;
    double K1, K2;        // D1 / D2 right-hand-side scale (stencil factor x solve scale)
    long N;               // points per field
    double *Q, *QW, *dQ;  // state, RK register, right-hand side (5 fields each)
    double *G;            // 12 first derivatives, G[(3*d + i)*N] = du_i/dx_d, G[(9+d)*N] = dT/dx_d
    double *L2e, *S, *Cb, *L2S, *L4e, *L4S;
    double *eta;          // 2 fields
    double *Hc;           // beta*, mu*, kappa*
    const double *dQp;    // z and x right-hand-side partials of the first-derivative passes (2 x 5 fields)
    const double *L2ep;   // z and x partials of the Laplacian of e (2 fields)
    const double *in0;    // single-operator input
    double *out0, *out1;  // single-operator outputs
    int iadd;             // single-operator accumulate flag
    int first;            // accumulators: 1 = overwrite
    int fix;              // z filter: apply the pressure fix
    int nslot;            // instances (slots) in the launch
    int dry;              // 1: check that the launch is valid, launch nothing
    int pad2_;
    double rk_a, rk_b, dt;
    // A batch stacks same-size instances field by field: point i of slot s
    // is element s*ni + i of each field of N = nslot*ni points.
    long ni;              // points per instance
    const int *slots;     // the launch's slots (device, nslot entries); null: 0..nslot-1
    const int *act;       // per-slot flag (device): 0 skips the slot; null: all
    const double *sre;    // per-slot 1/Re (device); null: ph.re_inv
    const double *sdt;    // per-slot dt (device); null: dt
    // sponge folded into the stage update: sigma per x index, then the left
    // and right reference states (5 each); null: no sponge in the update
    const double *spg;
    int spg_left;         // x indices below this take the left state
    int pad3_;
};

// The parameters of slot `slot` (its 1/Re and dt) over the launch-wide ones.
__device__ __forceinline__ void slot_params(const PassArgs &a, int slot, PassArgs &out)
{
    out = a;
    if (a.sre) out.ph.re_inv = a.sre[slot];
    if (a.sdt) out.dt = a.sdt[slot];
}

enum PassKind {
    K_D1, K_D2, K_FILT, K_GAUSS,
    K_GRAD,
    K_LAPSE,
    K_LAP2, K_LAP2_LAST,
    K_GAUSS_ETA, K_GAUSS_ETA_LAST,
    K_HDIV,
    K_RKF, K_FQ, K_FQ_FIX,
    K_GRAD_CROSS,
    K_HDIV_RKF,
    K_LAPSE_LAP2,
    K_NPASS,              // pass kinds end here
};

// One ragged launch: a pass over a group of batches of different grid sizes
// (cubic grids of 48, 64, 96 or 128 points, member c of the size family,
// see ragged_member).  Per member: its host tables (null: absent), grid and
// pass arguments; `first` is the member whose arguments pick the pass
// variant (all members run the same stage sequence).
struct RaggedSpec {
    int dir;
    int first;
    const DirTab *t[4];
    int My[4], Mz[4], Mx[4];
    PassArgs a[4];
};
// The size family of ragged launches: lines of RAGGED_N[c] points in the
// layout of RAGGED_M[c] points per lane on RAGGED_P[c] lanes (each size's
// narrow layout).
constexpr int RAGGED_N[4] = {48, 64, 96, 128};
constexpr int RAGGED_M[4] = {3, 4, 3, 4};
constexpr int RAGGED_P[4] = {16, 16, 32, 32};
// Family member of a line length (-1: not in the family).
int ragged_member(int n);
hipError_t run_pass_ragged(int kind, const RaggedSpec &sp, hipStream_t s);
hipError_t launch_grad_ragged(int kind, const RaggedSpec &sp, hipStream_t s);
hipError_t launch_lapse_ragged(int kind, const RaggedSpec &sp, hipStream_t s);
hipError_t launch_hyper_ragged(int kind, const RaggedSpec &sp, hipStream_t s);
hipError_t launch_hdivrk_ragged(int kind, const RaggedSpec &sp, hipStream_t s);
hipError_t launch_filter_ragged(int kind, const RaggedSpec &sp, hipStream_t s);

// Default launch layout of a pass kind: narrow = 1 uses the tables with half
// the points per lane (twice the lanes per line); nt = threads per workgroup.
void pass_default_cfg(int kind, int dir, int *narrow, int *nt);
hipError_t run_pass(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
// per translation unit (st_pass_*.hip)
hipError_t launch_ops(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
hipError_t launch_grad(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
hipError_t launch_lapse(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
hipError_t launch_hyper(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
hipError_t launch_hdivrk(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);
hipError_t launch_filter(int kind, const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s);

}  // namespace st

namespace st {
// which: 0 = CFL (2 maxima), 1 = dt limits (3 minima), 2 = diagnostics (7 sums);
// partials per slot and block: partial[(slot*nblocks + block)*NV + v]
hipError_t run_reduce(int which, const PassArgs &a, double *partial, int *nblocks, hipStream_t s);
int reduce_blocks(long ni);
hipError_t run_fill(double *p, long n, double v, hipStream_t s);
hipError_t run_fill_slots(double *p, int nfields, long N, long ni, int nslot, const int *slots, const int *act,
                          double v, hipStream_t s);
hipError_t run_fixp(const PassArgs &a, double *Qx, hipStream_t s);
hipError_t run_dt(const double *pc, const double *pt, int nb, int nslot, const int *active, const int *target,
                  const int *stepc, double cfl_cal, int hyper, int first_step, int *run, double *dt, double *dt_last,
                  hipStream_t s);
hipError_t run_clock(int nslot, const int *run, const double *dt, double *time, int *stepc, int bump, hipStream_t s);
hipError_t run_sponge(const PassArgs &a, const double *spg, int Mx, int nleft, hipStream_t s);
// pointwise steps of Compute_hyper and Compute_Grad outside a stage (one
// slot; the pointers carry its offset, fields N apart)
hipError_t run_hyp_prep(const double *Q, const double *G, long N, long ni, int want_hyp, double *W, hipStream_t s);
hipError_t run_hyp_cb(const double *W, long N, long ni, double *Cb, hipStream_t s);
hipError_t run_sum3(const double *z, const double *x, const double *y, long ni, int absval, double *out, hipStream_t s);
hipError_t run_hyp_eta(const double *Q, long N, long ni, const Phys &ph, const double *d4S, const double *d4e,
                       double *eta, hipStream_t s);
hipError_t run_hyp_coef(const double *g1, const double *g2, const double *Cb, long N, long ni, const Phys &ph,
                        double *Hc, hipStream_t s);
// max |dil| and sum |omega|^2 partials (2 per block)
int gradred_blocks(long ni);
hipError_t run_gradred(const double *G, long N, long ni, double *partial, hipStream_t s);
}  // namespace st
