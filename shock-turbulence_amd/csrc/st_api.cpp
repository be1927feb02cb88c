// Host side of the C-ABI (include/st_api.h): context, buffers, pass
// sequencing for RHS / Hyper_RHS / LDDRK4, reductions and profiling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/st_api.h"
#include "st_internal.h"
#include "st_kernels.h"

namespace {

constexpr int NKIND = 24;          // pass kinds + reductions + fills
constexpr int KIND_RED = 21, KIND_FILL = 22;
static_assert(st::K_NPASS <= KIND_RED, "pass kinds overlap the other timed kinds");

// Fields read + written per point by one launch of each kind (RMW counts
// twice); the algorithmic bytes of a launch are this times 8 B per point.
const char *kind_names[NKIND] = {
    "d1", "d2", "filter", "gauss", "grad", "lap_se",
    "lap2", "lap2_gauss", "gauss_eta", "gauss_hdiv", "hyper_div", "rk_filter", "filter_q",
    "filter_q_fix", "grad_cross", "hdiv_rkf", "lap_se_lap2", "", "", "", "", "reduce", "fill", ""};

struct Event {
    hipEvent_t a, b;
    int kind;             // pass kind * 3 + direction
};

// A pass recorded instead of launched (the members of a group, see
// group_step): its kind, direction, prepared arguments and host tables.
struct Rec {
    int kind, dir;
    st::PassArgs a;
    const st::DirTab *t;
};

}  // namespace

struct st_ctx {
    st_grid g;
    st_params p;
    int Mx, My, Mz;
    int ns;                         // instances (slots) stacked per field
    long ni;                        // points per instance
    long N;                         // field stride: ns * ni
    double sx1[3];
    st::Phys ph;
    st::Coef C;
    st::DirTab th[6];               // [dir] M points per lane, [3+dir] M/2 points per lane
    int cfg_narrow[NKIND][3], cfg_nt[NKIND][3];
    st::DirTab *td = nullptr;
    double *buf = nullptr;
    double *Q, *QW, *dQ, *G, *L2e, *S, *Cb, *L2S, *L4e, *L4S, *eta, *Hc, *T5, *tB, *tC;
    double *dQp, *L2ep;             // z and x partials of the first-derivative passes
    double *red = nullptr;          // reduction partials (device)
    double *red_h = nullptr;        // pinned host copy
    long red_n = 0;                 // doubles in red / red_h
    // per-slot state of a batch (device): active flag, run flag of the
    // current step, step target, driver step counter, 1/Re, dt, time
    int *s_act = nullptr, *s_run, *s_target, *s_step, *s_list;
    double *s_re = nullptr, *s_dt, *s_time, *s_dtlast;
    std::vector<int> act_h;
    int nact = 0;                   // active slots, listed in s_list (launch grids cover these only)
    std::vector<hipGraphExec_t> graphs;   // one batch step, captured per active-slot count
    double *spg = nullptr;          // sponge: sigma per x index, then the left and right reference states
    int spg_left = 0;               // x indices [0, spg_left) take the left reference state
    int use_graph = 1;
    hipStream_t s = nullptr;
    char err[512];
    int prof = 0;
    std::vector<Event> ev;
    std::vector<hipEvent_t> pool;
    double prof_ms[NKIND * 3];      // per pass kind and direction
    long prof_cnt[NKIND * 3];
    double prof_fields[NKIND * 3];
    std::vector<Rec> *rec = nullptr;  // set: pass() records instead of launching
    int force_narrow = 0;             // 1: every pass takes the M/2-point layout
};

namespace {

int fail(st_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(c, x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) return fail(c, -3, "%s: %s", #x, hipGetErrorString(e_));       \
    } while (0)

hipEvent_t ev_get(st_ctx *c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(st_ctx *c, int kind, Event &e)
{
    e.kind = kind;
    e.a = ev_get(c);
    e.b = ev_get(c);
    (void)hipEventRecord(e.a, c->s);
}

void prof_end(st_ctx *c, Event &e)
{
    (void)hipEventRecord(e.b, c->s);
    c->ev.push_back(e);
}

// Collect finished events (after a synchronisation).
void prof_collect(st_ctx *c)
{
    for (auto &e : c->ev) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            c->prof_ms[e.kind] += ms;
            c->prof_cnt[e.kind] += 1;
        }
        c->pool.push_back(e.a);
        c->pool.push_back(e.b);
    }
    c->ev.clear();
}

st::PassArgs base_args(st_ctx *c)
{
    st::PassArgs a;
    std::memset(&a, 0, sizeof(a));
    a.C = c->C;
    a.ph = c->ph;
    a.N = c->N;
    a.ni = c->ni;
    a.nslot = 1;                    // single-instance entry points act on slot 0
    a.Q = c->Q;
    a.QW = c->QW;
    a.dQ = c->dQ;
    a.G = c->G;
    a.L2e = c->p.hyper_param ? c->L2e : nullptr;
    a.S = c->S;
    a.Cb = c->g.nonperiodic_x ? nullptr : c->Cb;
    a.L2S = c->L2S;
    a.L4e = c->L4e;
    a.L4S = c->L4S;
    a.eta = c->eta;
    a.Hc = c->Hc;
    a.dQp = c->dQp;
    a.L2ep = c->L2ep;
    return a;
}

// fields read+written per point for one launch of a pass kind
double kind_fields(int kind, int dir, const st::PassArgs &a)
{
    switch (kind) {
    case st::K_D1: return a.iadd ? 3 : 2;
    case st::K_D2: return a.iadd ? 3 : 2;
    case st::K_FILT: return 2;
    case st::K_GAUSS: return 2;
    // Q, G of the direction, the dQ partial, the Lap e partial; z also reads
    // the two stored derivatives of its mixed terms
    case st::K_GRAD: return 5 + 4 + 5 + (a.L2e ? 1 : 0) + (dir == 2 ? 2 : 0);
    case st::K_LAPSE: return a.first ? 2 + 2 : 2 + 4;
    case st::K_LAP2: return a.first ? 2 : 3;
    // L2S, L4S, L4e, Q in; eta out
    case st::K_LAP2_LAST: return 1 + 1 + 1 + 5 + 2;
    // S, Lap e and the L2S, L4e partials in; L2S, L4e, L4S out
    case st::K_LAPSE_LAP2: return 4 + 3;
    case st::K_GAUSS_ETA: return 4;
    // eta, C_beta, beta* mu* kappa* (written), the z-flux inputs (8 G, 4 Q), dQ 4 (read and written)
    case st::K_GAUSS_ETA_LAST: return 2 + (a.Cb ? 1 : 0) + 3 + 8 + 4 + 8;
    case st::K_HDIV: return 3 + 8 + 4 + 8;
    case st::K_RKF: return 15 + 5 + 5;
    case st::K_FQ: return 10;
    case st::K_FQ_FIX: return 10;
    // the divergence inputs (Hc, 8 G, Q), dQ, QW; QW and Q written
    case st::K_HDIV_RKF: return 3 + 8 + 5 + 5 + 5 + 5 + 5;
    // Q, G of y (written), 8 G of x and z (the four inputs of the mixed
    // terms and four more), the z and x partials, dQ (written); with the
    // artificial transport the Laplacian of e (2 partials in, the sum out),
    // |S| and C_beta
    case st::K_GRAD_CROSS: return 5 + 4 + 8 + 10 + 5 + (a.L2e ? 3 + 1 + (a.Cb ? 1 : 0) : 0);
    default: return 0;
    }
}

int pass(st_ctx *c, int kind, int dir, st::PassArgs a, double cadd = 1.0)
{
    const int pk = kind * 3 + dir;
    const int lay = c->force_narrow || c->cfg_narrow[kind][dir] ? 3 : 0;
    const st::DirTab &t = c->th[lay + dir];
    a.tab = c->td + lay + dir;
    a.K1 = cadd * c->sx1[dir] * (double)(t.n - 1) * t.op[0].scale;   // fac of src/derivCDopt6_periodic.f:181
    a.K2 = t.s2 * t.op[1].scale;
    if (c->rec) {
        c->rec->push_back(Rec{kind, dir, a, &t});
        return 0;
    }
    Event e;
    const int prof = c->prof && !a.dry;
    if (prof) prof_begin(c, pk, e);
    // the widest configured tile that the line count and layout allow
    int nt = c->cfg_nt[kind][dir];
    const long lot = c->ni / t.n;
    // (1024-thread tiles exist for M = 4, 8 with 32 or 64 lanes per line)
    while (nt > 256 && ((nt == 1024 && (t.P < 32 || (t.M & (t.M - 1)))) || lot % (nt / t.P))) nt /= 2;
    hipError_t err = st::run_pass(kind, t, nt, dir, c->My, c->Mz, c->Mx, a, c->s);
    if (a.dry) return err == hipSuccess ? 0 : -4;
    // A shape this grid cannot take (the tile's LDS, say) is refused before
    // any launch: the next narrower tiles of the same layout, then those of
    // the other layout.  The defaults are tuned at 256^3 and need not fit
    // every line length.
    for (int alt = 0; err == hipErrorInvalidValue && alt < 2; alt++) {
        const int la = alt ? 3 - lay : lay;
        const st::DirTab &u = c->th[la + dir];
        st::PassArgs b = a;
        b.tab = c->td + la + dir;
        b.K1 = cadd * c->sx1[dir] * (double)(u.n - 1) * u.op[0].scale;
        b.K2 = u.s2 * u.op[1].scale;
        for (int w = alt ? 1024 : nt / 2; w >= 256 && err == hipErrorInvalidValue; w /= 2) {
            if ((w == 1024 && (u.P < 32 || (u.M & (u.M - 1)))) || lot % (w / u.P)) continue;
            err = st::run_pass(kind, u, w, dir, c->My, c->Mz, c->Mx, b, c->s);
        }
    }
    if (prof) {
        prof_end(c, e);
        c->prof_fields[pk] = kind_fields(kind, dir, a);
    }
    if (err != hipSuccess) return fail(c, -4, "pass %s dir %d: %s", kind_names[kind], dir, hipGetErrorString(err));
    return 0;
}

}  // namespace


// ---------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------
extern "C" int st_batch_create(const st_grid *grid, const st_params *params, int nslots, st_ctx **out)
{
    if (!grid || !params || !out || nslots < 1) return -1;
    *out = nullptr;
    st_ctx *c = new st_ctx();
    c->err[0] = 0;
    c->g = *grid;
    c->p = *params;
    c->Mx = grid->nx;
    c->My = grid->ny;
    c->Mz = grid->nz;
    c->ns = nslots;
    c->ni = (long)c->Mx * c->My * c->Mz;
    c->N = c->ni * nslots;
    c->act_h.assign(nslots, 0);
    c->graphs.assign(nslots + 1, nullptr);
    std::memset(c->prof_ms, 0, sizeof(c->prof_ms));
    std::memset(c->prof_cnt, 0, sizeof(c->prof_cnt));
    std::memset(c->prof_fields, 0, sizeof(c->prof_fields));
    // launch layouts; ST_NARROW / ST_NT512 / ST_NT1024 (bit masks over pass
    // kinds) and ST_NT_<kind> (per direction) override the defaults
    {
        const char *en = getenv("ST_NARROW"), *et = getenv("ST_NT512"), *eu = getenv("ST_NT1024");
        for (int k = 0; k < NKIND; k++) {
            for (int d = 0; d < 3; d++) {
                st::pass_default_cfg(k, d, &c->cfg_narrow[k][d], &c->cfg_nt[k][d]);
                if (et) c->cfg_nt[k][d] = ((strtoul(et, nullptr, 0) >> k) & 1) ? 512 : 256;
                if (eu && ((strtoul(eu, nullptr, 0) >> k) & 1)) c->cfg_nt[k][d] = 1024;
            }
            if (en)
                for (int d = 0; d < 3; d++) c->cfg_narrow[k][d] = (int)((strtoul(en, nullptr, 0) >> k) & 1);
            // ST_NT_<k> = "nx,ny,nz": tile threads of pass kind k per direction
            char var[32];
            snprintf(var, sizeof(var), "ST_NT_%d", k);
            if (const char *ev = getenv(var)) {
                int v[3];
                if (sscanf(ev, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
                    for (int d = 0; d < 3; d++)
                        if (v[d] == 256 || v[d] == 512 || v[d] == 1024) c->cfg_nt[k][d] = v[d];
            }
        }
    }
    if (grid->nonperiodic_x && !getenv("ST_NT1024"))
        for (int k = 0; k < NKIND; k++) c->cfg_nt[k][0] = 256;   // boundary tables cost registers
    const int n[3] = {c->Mx, c->My, c->Mz};
    const double period[3] = {grid->period_x, grid->period_y, grid->period_z};
    const double start[3] = {grid->x_start, grid->y_start, grid->z_start};
    for (int d = 0; d < 3; d++) {
        int M, P;
        if (n[d] < 16 || !st::choose_layout(n[d], &M, &P)) {
            fail(c, -2, "unsupported line length %d in direction %d (need n = M*P, M in {2,4,5,6,7,8}, P in "
                        "{8,16,32,64}: 16..512 points, a power of two or 3, 5 or 7 times one)", n[d], d);
            *out = c;
            return -2;
        }
        const int B = 256 / P;             // lines per tile (same for the 512-thread layout)
        const long lot = c->ni / n[d];
        if (lot % B) {
            fail(c, -2, "line count %ld of direction %d not a multiple of the tile", lot, d);
            *out = c;
            return -2;
        }
    }
    // physical constants
    const double gamma = params->gamma;
    c->ph.gamma = gamma;
    c->ph.gm1 = gamma - 1.0;
    c->ph.gfac = gamma / (gamma - 1.0);
    c->ph.gm1_inv = 1.0 / (gamma - 1.0);
    c->ph.ginv = 1.0 / gamma;
    c->ph.re_inv = params->Re > 0.0 ? 1.0 / params->Re : 0.0;
    c->ph.pr_inv = params->Pr > 0.0 ? 1.0 / params->Pr : 0.0;
    c->ph.vis_exp = params->vis_exp;
    c->ph.vis_exp_m1 = params->vis_exp - 1.0;
    c->ph.bulk = params->bulk_ratio;
    c->ph.cmu = 2.0e-3;
    c->ph.ck = 1.0e-2;
    c->ph.cfl_cal = params->cfl_cal;
    const int np[3] = {grid->nonperiodic_x ? 1 : 0, 0, 0};
    for (int d = 0; d < 3; d++) c->ph.D[d] = st::grid_spacing(period[d], start[d], n[d], np[d]);   // src/grid.f:40
    c->ph.dx6 = std::pow(c->ph.D[0], 6.0);
    c->ph.dx5 = std::pow(c->ph.D[0], 5.0);
    // scheme coefficients (src/header:193-203, src/LES_header:107-113,153-157)
    st::Coef &C = c->C;
    C.a1 = 14.0 / 18.0;
    C.b1 = 1.0 / 36.0;
    C.c1 = 0.0;
    C.a2 = 12.0 / 11.0;
    C.b2 = 3.0 / 44.0;
    if (params->d2_mode == 1) {
        C.c2 = 2.0 * (C.a2 + C.b2);
        C.d2 = 0.0;
    } else {
        C.c2 = 0.0;
        C.d2 = 2.0 * (C.a2 + C.b2);
    }
    C.f0 = 9.9965e-01;
    C.f1 = 0.66652e+00;
    C.f2 = 0.16674e+00;
    C.f3 = 4.0e-05;
    C.f4 = -5.0e-06;
    C.g0 = 3565.0 / 10368.0;
    C.g1 = 3091.0 / 12960.0;
    C.g2 = 1997.0 / 25920.0;
    C.g3 = 149.0 / 12960.0;
    C.g4 = 107.0 / 103680.0;
    for (int d = 0; d < 3; d++) {
        int M, P;
        st::choose_layout(n[d], &M, &P);
        const double sx1 = 1.0 / st::grid_leng(period[d], n[d], np[d]);   // src/grid.f:31-32
        const double sx2 = sx1 * sx1;
        c->sx1[d] = sx1;
        // the narrow layout (M/2 points per lane, 2P lanes) where it exists;
        // 512-point lines already use 64 lanes and keep M = 8
        const bool narrow = M % 2 == 0 && 2 * P <= 64;
        if (st::build_dirtab(&c->th[d], n[d], M, P, 0, sx1, sx2, 1.0 / 3.0, 2.0 / 11.0, 6.6624e-01, 1.6688e-01,
                             np[d]) ||
            (narrow && st::build_dirtab(&c->th[3 + d], n[d], M / 2, 2 * P, 0, sx1, sx2, 1.0 / 3.0, 2.0 / 11.0,
                                        6.6624e-01, 1.6688e-01, np[d]))) {
            fail(c, -2, "coefficient tables failed for direction %d", d);
            *out = c;
            return -2;
        }
        if (!narrow) c->th[3 + d] = c->th[d];
    }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking)) != hipSuccess) {
        fail(c, -3, "hipStreamCreate: %s", hipGetErrorString(e));
        *out = c;
        return -3;
    }
    const long nf = 5 + 5 + 5 + 12 + 1 + 1 + 1 + 1 + 1 + 1 + 2 + 3 + 5 + 1 + 1 + 10 + 2;
    // reduction partials: up to 7 values per block and slot, for the CFL
    // and dt-limit folds side by side (2 + 3 values)
    c->red_n = (long)st::reduce_blocks(c->ni) * nslots * 8;
    void *sp = nullptr;
    const size_t ns = (size_t)nslots;
    if ((e = hipMalloc(&c->buf, sizeof(double) * nf * c->N)) != hipSuccess ||
        (e = hipMalloc(&c->td, sizeof(st::DirTab) * 6)) != hipSuccess ||
        (e = hipMalloc(&c->red, sizeof(double) * c->red_n)) != hipSuccess ||
        (e = hipHostMalloc(&c->red_h, sizeof(double) * c->red_n)) != hipSuccess ||
        (e = hipMalloc(&sp, ns * (6 * sizeof(int) + 4 * sizeof(double)))) != hipSuccess) {
        fail(c, -3, "device allocation: %s", hipGetErrorString(e));
        *out = c;
        return -3;
    }
    c->s_re = (double *)sp;
    c->s_dt = c->s_re + ns;
    c->s_time = c->s_dt + ns;
    c->s_dtlast = c->s_time + ns;
    c->s_act = (int *)(c->s_dtlast + ns);
    c->s_run = c->s_act + ns;
    c->s_target = c->s_run + ns;
    c->s_step = c->s_target + ns;
    c->s_list = c->s_step + ns;
    double *p = c->buf;
    const long N = c->N;
    c->Q = p; p += 5 * N;
    c->QW = p; p += 5 * N;
    c->dQ = p; p += 5 * N;
    c->G = p; p += 12 * N;
    c->L2e = p; p += N;
    c->S = p; p += N;
    c->Cb = p; p += N;
    c->L2S = p; p += N;
    c->L4e = p; p += N;
    c->L4S = p; p += N;
    c->eta = p; p += 2 * N;
    c->Hc = p; p += 3 * N;
    c->T5 = p; p += 5 * N;
    c->tB = p; p += N;
    c->tC = p; p += N;
    c->dQp = p; p += 10 * N;
    c->L2ep = p; p += 2 * N;
    if ((e = hipMemcpyAsync(c->td, c->th, sizeof(st::DirTab) * 6, hipMemcpyHostToDevice, c->s)) != hipSuccess ||
        (e = st::run_fill(c->buf, nf * N, 0.0, c->s)) != hipSuccess ||
        (e = st::run_fill(c->s_re, 4 * (long)ns + 3 * (long)ns, 0.0, c->s)) != hipSuccess ||
        (e = hipStreamSynchronize(c->s)) != hipSuccess) {
        fail(c, -3, "device init: %s", hipGetErrorString(e));
        *out = c;
        return -3;
    }
    *out = c;
    return 0;
}

extern "C" int st_ctx_create(const st_grid *grid, const st_params *params, st_ctx **out)
{
    return st_batch_create(grid, params, 1, out);
}

extern "C" void st_ctx_destroy(st_ctx *c)
{
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (auto &e : c->ev) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    for (auto g : c->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->s_re) (void)hipFree(c->s_re);
    if (c->spg) (void)hipFree(c->spg);
    if (c->buf) (void)hipFree(c->buf);
    if (c->td) (void)hipFree(c->td);
    if (c->red) (void)hipFree(c->red);
    if (c->red_h) (void)hipHostFree(c->red_h);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

namespace {
thread_local char call_err[512];          // last failure of a call without a context
}

int st::set_call_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(call_err, sizeof(call_err), fmt, ap);
    va_end(ap);
    return code;
}

#ifndef ST_SRC_HASH
#define ST_SRC_HASH "unknown"
#endif

extern "C" const char *st_build_id(void) { return "src-sha256:" ST_SRC_HASH; }

extern "C" const char *st_last_error(const st_ctx *c)
{
    if (c) return c->err;
    return call_err[0] ? call_err : "null context";
}

extern "C" int st_sync(st_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    return 0;
}

// ---------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------
namespace {

// RK coefficient tables: CC5 (src/header:284-296), LDDRK6 (src/header:267-281,
// identical to init/header:222-236), LDDRK5 (src/header:253-265).
const double rk_cc5_a[5] = {0.0, -4.801594388477952e-01, -1.404247195200064e+00, -2.016477077503357e+00, -1.056444269767207e+00};
const double rk_cc5_b[5] = {1.028639988104959e-01, 7.408540575766301e-01, 7.426530946683799e-01, 4.694937902357905e-01, 1.881733382887932e-01};
const double rk_ld6_a[6] = {0.0, -4.4127377153877382565e-1, -1.073982008079781868e0, -1.7063570791256758809e0,
                            -2.7979293162682443056e0, -4.0913537120919160454e0};
const double rk_ld6_b[6] = {1.1584888181285561688e-1, 3.7287699051652864918e-1, 7.3795368921435295698e-1,
                            5.7981109366311039538e-1, 1.031284991300145194e0, 1.5e-1};
const double rk_ld5_a[5] = {0.0, -6.0512264332862261228e-1, -2.04375640234761394333e0, -0.74069990637544192841e0,
                            -4.42317651302968168941e0};
const double rk_ld5_b[5] = {2.6874543888713438496e-1, 8.014706973220802933e-1, 5.0515704269422722538e-1,
                            5.6235680379000296407e-1, 5.900655127758823335e-2};

int rk_loop(int scheme, int loop, const double **a, const double **b)
{
    if (loop == 0) {
        if (scheme == 1) { *a = rk_ld5_a; *b = rk_ld5_b; return 5; }
        *a = rk_cc5_a; *b = rk_cc5_b; return 5;
    }
    if (loop == 1 && scheme != 2) { *a = rk_ld6_a; *b = rk_ld6_b; return 6; }
    return 0;
}

// Host arrays hold one instance (nfields x ni, contiguous); on the device
// the fields of slot `slot` sit at dev + f*N + slot*ni.
int up(st_ctx *c, double *dev, const double *host, long nfields, int slot = 0)
{
    const size_t w = sizeof(double) * c->ni;
    HIPCHK(c, hipMemcpy2DAsync(dev + (size_t)slot * c->ni, sizeof(double) * c->N, host, w, w, nfields,
                               hipMemcpyHostToDevice, c->s));
    return 0;
}

int down(st_ctx *c, double *host, const double *dev, long nfields, int slot = 0)
{
    const size_t w = sizeof(double) * c->ni;
    HIPCHK(c, hipMemcpy2DAsync(host, w, dev + (size_t)slot * c->ni, sizeof(double) * c->N, w, nfields,
                               hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    return 0;
}

// fields [p, p + nfields*N) of the launch's slots (all, or those with act)
int fill(st_ctx *c, double *p, int nfields, const st::PassArgs &base, double v)
{
    Event e;
    if (c->prof) prof_begin(c, KIND_FILL * 3, e);
    hipError_t err = st::run_fill_slots(p, nfields, c->N, c->ni, base.nslot, base.slots, base.act, v, c->s);
    if (c->prof) prof_end(c, e);
    if (err != hipSuccess) return fail(c, -4, "fill: %s", hipGetErrorString(err));
    return 0;
}

#define TRY(x)                \
    do {                      \
        int r_ = (x);         \
        if (r_) return r_;    \
    } while (0)

// RHS (src/rhsv_periodic_sf.f:9-25) into dQ; with the artificial transport
// on, also the strain magnitude, the sensor and the x and y parts of the
// first Laplacians it needs.
int seq_rhs(st_ctx *c, const st::PassArgs &base)
{
    const int hyp = c->p.hyper_param;
    st::PassArgs a = base;
    a.first = 1;
    // the x and z passes write their own partial sums (no read-modify-write);
    // the y pass adds them up with its own terms and the viscous products.
    // z and y also take the mixed second derivatives, from the x (and z)
    // derivatives stored before them.
    {
        st::PassArgs b = a;
        b.dQ = c->dQp + 5 * c->N;
        b.L2e = hyp ? c->L2ep + c->N : nullptr;
        TRY(pass(c, st::K_GRAD, 0, b));
        b.dQ = c->dQp;
        b.L2e = hyp ? c->L2ep : nullptr;
        TRY(pass(c, st::K_GRAD, 2, b));
    }
    TRY(pass(c, st::K_GRAD_CROSS, 1, a));
    if (hyp) {
        // D2 of |S| and Lap e, summed over x, y (z joins the first pass of
        // the artificial transport)
        TRY(pass(c, st::K_LAPSE, 0, a));
        a.first = 0;
        TRY(pass(c, st::K_LAPSE, 1, a));
    }
    return 0;
}

// Hyper_RHS (src/hyper_rhs.f:12-499) added into dQ, directions z, x, y.
// With rk set, the last (y) divergence pass also applies the stage update
// and the y filter (K_HDIV_RKF), so the stage goes on with the z and x
// filters.
int seq_hyper(st_ctx *c, const st::PassArgs &base, const st::PassArgs *rk = nullptr)
{
    if (!c->p.hyper_param) return 0;
    st::PassArgs a = base;
    a.first = 1;
    // z: the last D2 of |S| and Lap e, then the first of Lap(Lap|S|)
    TRY(pass(c, st::K_LAPSE_LAP2, 2, a));
    a.first = 0;
    TRY(pass(c, st::K_LAP2, 0, a));
    // y: the last of Lap(Lap|S|), eta, and the first Gaussian direction
    TRY(pass(c, st::K_LAP2_LAST, 1, a));
    TRY(pass(c, st::K_GAUSS_ETA, 0, a));
    TRY(pass(c, st::K_GAUSS_ETA_LAST, 2, a));            // and the z divergence
    TRY(pass(c, st::K_HDIV, 0, a));
    if (rk) return pass(c, st::K_HDIV_RKF, 1, *rk);
    TRY(pass(c, st::K_HDIV, 1, a));
    return 0;
}

// One stage body (src/lddrk4_periodic_sf.f:344-356).
int seq_stage(st_ctx *c, const st::PassArgs &base, double ra, double rb, double dt)
{
    st::PassArgs a = base;
    a.rk_a = ra;
    a.rk_b = rb;
    a.dt = dt;
    TRY(seq_rhs(c, base));
    // the update (and the sponge term before it) joins the last divergence
    // pass
    const bool fuse = c->p.hyper_param && !getenv("ST_NO_FUSE_RK");
    if (fuse && c->spg) {
        a.spg = c->spg;
        a.spg_left = c->spg_left;
    }
    TRY(seq_hyper(c, base, fuse ? &a : nullptr));
    if (c->spg && !fuse) {                                   // the hook order of src/lddrk4_nonperiodic_mf.f:462-466
        const hipError_t e = st::run_sponge(base, c->spg, c->Mx, c->spg_left, c->s);
        if (e != hipSuccess) return fail(c, -4, "sponge: %s", hipGetErrorString(e));
    }
    // filters y (with the update), z, x (with Fix_Pressure)
    if (!fuse) TRY(pass(c, st::K_RKF, 1, a));
    TRY(pass(c, st::K_FQ, 2, a));
    a.fix = 1;
    TRY(pass(c, st::K_FQ_FIX, 0, a));
    return 0;
}

int reduce(st_ctx *c, int which, int nv, double *out)
{
    st::PassArgs a = base_args(c);
    int nb = 0;
    Event e;
    if (c->prof) prof_begin(c, KIND_RED * 3, e);
    hipError_t err = st::run_reduce(which, a, c->red, &nb, c->s);
    if (c->prof) prof_end(c, e);
    if (err != hipSuccess) return fail(c, -4, "reduce: %s", hipGetErrorString(err));
    HIPCHK(c, hipMemcpyAsync(c->red_h, c->red, sizeof(double) * nb * nv, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    for (int i = 0; i < nv; i++) {
        double v = c->red_h[i];
        for (int b = 1; b < nb; b++) {
            const double x = c->red_h[b * nv + i];
            if (which == 0) v = x > v ? x : v;
            else if (which == 1) v = x < v ? x : v;
            else v += x;
        }
        out[i] = v;
    }
    return 0;
}

int kerr(st_ctx *c, hipError_t e, const char *what)
{
    if (e != hipSuccess) return fail(c, -4, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

// One single-operator pass over one slot's fields (the pointers carry the
// slot's offset; the launch covers slot 0 of a one-slot argument set).
int op1(st_ctx *c, int kind, int dir, const double *in, double *out0, double *out1, int iadd, double cadd)
{
    st::PassArgs a = base_args(c);
    a.in0 = in;
    a.out0 = out0;
    a.out1 = out1;
    a.iadd = iadd;
    return pass(c, kind, dir, a, cadd);
}

// Compute_Laplacian (src/hyper_dynamic.f:41-58): out = D2z + D2x + D2y of f
// (|.| with absval, :278-279); tmp holds 3 fields.
int laplacian(st_ctx *c, const double *f, double *out, double *tmp, int absval)
{
    const long N = c->N;
    TRY(op1(c, st::K_D2, 0, f, nullptr, tmp, 0, 1.0));
    TRY(op1(c, st::K_D2, 1, f, nullptr, tmp + N, 0, 1.0));
    TRY(op1(c, st::K_D2, 2, f, nullptr, tmp + 2 * N, 0, 1.0));
    return kerr(c, st::run_sum3(tmp + 2 * N, tmp, tmp + N, c->ni, absval, out, c->s), "laplacian sum");
}

// Compute_hyper (src/hyper_dynamic.f:173-318; with a non-periodic x the
// static coefficients of src/hyper_routines.f:66-214) for one slot, outside
// a stage: beta*, mu*, kappa* into Hc from the resident state and, with
// use_grad, the velocity gradients in G (/GRAD/ as the last RHS left it);
// without, |S| = 0, as at a start before any RHS (src/main.f:100 with
// /GRAD/ still zero).  Scratch: the 10 partial-sum fields and T5, which a
// stage rewrites before reading.
int seq_compute_hyper(st_ctx *c, int slot, int use_grad)
{
    const long off = (long)slot * c->ni, N = c->N, ni = c->ni;
    const double *Q = c->Q + off;
    double *W = c->dQp + off;
    double *T = c->T5 + off;
    TRY(kerr(c, st::run_hyp_prep(Q, use_grad ? c->G + off : nullptr, N, ni, 1, W, c->s), "hyper prep"));
    double *Cb = nullptr;
    if (!c->g.nonperiodic_x) {
        // Compute_coeff1 (:72-165): vorticity W[3..5] and dilatation W[6] of U
        double *U0 = W, *U1 = W + N, *U2 = W + 2 * N;
        double *wx = W + 3 * N, *wy = W + 4 * N, *wz = W + 5 * N, *dil = W + 6 * N;
        TRY(op1(c, st::K_D1, 1, U2, wx, nullptr, 0, 1.0));            // :105-106
        TRY(op1(c, st::K_D1, 2, U1, wx, nullptr, 1, -1.0));
        TRY(op1(c, st::K_D1, 0, U2, wy, nullptr, 0, -1.0));           // :114-115
        TRY(op1(c, st::K_D1, 2, U0, wy, nullptr, 1, 1.0));
        TRY(op1(c, st::K_D1, 0, U1, wz, nullptr, 0, 1.0));            // :123-124
        TRY(op1(c, st::K_D1, 1, U0, wz, nullptr, 1, -1.0));
        TRY(op1(c, st::K_D1, 0, U0, dil, nullptr, 0, 1.0));           // :132-140
        TRY(op1(c, st::K_D1, 1, U1, dil, nullptr, 1, 1.0));
        TRY(op1(c, st::K_D1, 2, U2, dil, nullptr, 1, 1.0));
        Cb = c->Cb + off;
        TRY(kerr(c, st::run_hyp_cb(W, N, ni, Cb, c->s), "C_beta"));
    }
    // |Lap^2 S| into W[7], |Lap^2 e| into W[8] (:272-279)
    TRY(laplacian(c, W + 7 * N, W + 9 * N, T, 0));
    TRY(laplacian(c, W + 9 * N, W + 7 * N, T, 1));
    TRY(laplacian(c, W + 8 * N, W + 9 * N, T, 0));
    TRY(laplacian(c, W + 9 * N, W + 8 * N, T, 1));
    // eta (:284-297), Gaussian x, y, z (:300), the coefficients (:303-314)
    double *eta = c->eta + off;
    TRY(kerr(c, st::run_hyp_eta(Q, N, ni, c->ph, W + 7 * N, W + 8 * N, eta, c->s), "eta"));
    TRY(op1(c, st::K_GAUSS, 0, eta, T, nullptr, 0, 1.0));
    TRY(op1(c, st::K_GAUSS, 1, T, T + N, nullptr, 0, 1.0));
    TRY(op1(c, st::K_GAUSS, 2, T + N, W + 9 * N, nullptr, 0, 1.0));
    TRY(op1(c, st::K_GAUSS, 0, eta + N, T + 2 * N, nullptr, 0, 1.0));
    TRY(op1(c, st::K_GAUSS, 1, T + 2 * N, T + 3 * N, nullptr, 0, 1.0));
    TRY(op1(c, st::K_GAUSS, 2, T + 3 * N, W + 7 * N, nullptr, 0, 1.0));
    return kerr(c, st::run_hyp_coef(W + 9 * N, W + 7 * N, Cb, N, ni, c->ph, c->Hc + off, c->s), "coefficients");
}

int stage_loop(st_ctx *c, double dt, double *time, int *step)
{
    const st::PassArgs base = base_args(c);
    for (int loop = 0; loop < 2; loop++) {
        const double *ra, *rb;
        const int ns = rk_loop(c->p.rk_scheme, loop, &ra, &rb);
        if (!ns) break;
        // QW = 0 at the loop start (:339, :376): the first stage's alpha is 0
        // and the update then does not read QW
        for (int k = 0; k < ns; k++) TRY(seq_stage(c, base, ra[k], rb[k], dt));
        if (time) *time = *time + dt;                        // :372, :419
    }
    if (step) *step = *step + 1;                             // :420
    return 0;
}

// One step of every running slot of a batch, entirely on the device: CFL
// and dt-limit partials, dt per slot (k_dt), then the two RK loops with
// each slot's own dt and 1/Re, and the slot clocks.  No host round trip,
// so a step can be captured in a graph and replayed.
int batch_step(st_ctx *c)
{
    if (!c->nact) return 0;
    st::PassArgs a = base_args(c);
    a.nslot = c->nact;
    a.slots = c->s_list;
    a.act = c->s_act;
    int nb = 0;
    double *pc = c->red, *pt = c->red + (long)st::reduce_blocks(c->ni) * c->ns * 2;
    hipError_t err = st::run_reduce(0, a, pc, &nb, c->s);
    if (err == hipSuccess && c->p.hyper_param == 1) err = st::run_reduce(1, a, pt, &nb, c->s);
    if (err == hipSuccess)
        err = st::run_dt(pc, pt, nb, c->ns, c->s_act, c->s_target, c->s_step, c->p.cfl_cal, c->p.hyper_param,
                         c->g.nonperiodic_x ? 1 : 0, c->s_run, c->s_dt, c->s_dtlast, c->s);
    if (err != hipSuccess) return fail(c, -4, "batch dt: %s", hipGetErrorString(err));
    a.act = c->s_run;
    a.sre = c->s_re;
    a.sdt = c->s_dt;
    for (int loop = 0; loop < 2; loop++) {
        const double *ra, *rb;
        const int ns = rk_loop(c->p.rk_scheme, loop, &ra, &rb);
        if (!ns) break;
        for (int k = 0; k < ns; k++) TRY(seq_stage(c, a, ra[k], rb[k], 0.0));   // (QW: as stage_loop)
        const int last = loop == 1 || !rk_loop(c->p.rk_scheme, 1, &ra, &rb);
        err = st::run_clock(c->ns, c->s_run, c->s_dt, c->s_time, c->s_step, last, c->s);
        if (err != hipSuccess) return fail(c, -4, "batch clock: %s", hipGetErrorString(err));
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------
extern "C" int st_state_upload(st_ctx *c, const double *Q)
{
    if (!c || !Q) return -1;
    TRY(up(c, c->Q, Q, 5));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int st_state_download(st_ctx *c, double *Q)
{
    if (!c || !Q) return -1;
    return down(c, Q, c->Q, 5);
}

extern "C" int st_state_dev(st_ctx *c, double **Q_dev)
{
    if (!c || !Q_dev) return -1;
    *Q_dev = c->Q;
    return 0;
}

extern "C" int st_d1(st_ctx *c, int dir, const double *F, double *dF, int iadd, double cadd)
{
    if (!c || !F || !dF || dir < 0 || dir > 2) return fail(c, -1, "st_d1: bad arguments");
    TRY(up(c, c->T5, F, 1));
    if (iadd) TRY(up(c, c->tB, dF, 1));
    st::PassArgs a = base_args(c);
    a.in0 = c->T5;
    a.out0 = c->tB;
    a.iadd = iadd ? 1 : 0;
    TRY(pass(c, st::K_D1, dir, a, cadd));
    return down(c, dF, c->tB, 1);
}

extern "C" int st_d2(st_ctx *c, int dir, const double *F, double *dF, double *d2F, int iadd)
{
    if (!c || !F || !d2F || dir < 0 || dir > 2 || (iadd && !dF)) return fail(c, -1, "st_d2: bad arguments");
    TRY(up(c, c->T5, F, 1));
    st::PassArgs a = base_args(c);
    a.in0 = c->T5;
    a.out0 = c->tB;
    a.out1 = c->tC;
    a.iadd = iadd ? 1 : 0;
    TRY(pass(c, st::K_D2, dir, a));
    if (iadd) TRY(down(c, dF, c->tB, 1));
    return down(c, d2F, c->tC, 1);
}

extern "C" int st_filter_dir(st_ctx *c, int dir, const double *F, double *F0)
{
    if (!c || !F || !F0 || dir < 0 || dir > 2) return fail(c, -1, "st_filter_dir: bad arguments");
    TRY(up(c, c->T5, F, 1));
    st::PassArgs a = base_args(c);
    a.in0 = c->T5;
    a.out0 = c->tB;
    TRY(pass(c, st::K_FILT, dir, a));
    return down(c, F0, c->tB, 1);
}

extern "C" int st_gauss_dir(st_ctx *c, int dir, const double *F, double *Fg)
{
    if (!c || !F || !Fg || dir < 0 || dir > 2) return fail(c, -1, "st_gauss_dir: bad arguments");
    TRY(up(c, c->T5, F, 1));
    st::PassArgs a = base_args(c);
    a.in0 = c->T5;
    a.out0 = c->tB;
    TRY(pass(c, st::K_GAUSS, dir, a));
    return down(c, Fg, c->tB, 1);
}

static int chain3(st_ctx *c, int kind, const double *F, double *F0, int nf)
{
    for (int L = 0; L < nf; L++) {
        TRY(up(c, c->tB, F + (size_t)L * c->ni, 1));
        st::PassArgs a = base_args(c);
        a.in0 = c->tB;
        a.out0 = c->tC;
        TRY(pass(c, kind, 0, a));
        a.in0 = c->tC;
        a.out0 = c->tB;
        TRY(pass(c, kind, 1, a));
        a.in0 = c->tB;
        a.out0 = c->tC;
        TRY(pass(c, kind, 2, a));
        TRY(down(c, F0 + (size_t)L * c->ni, c->tC, 1));
    }
    return 0;
}

extern "C" int st_filter(st_ctx *c, const double *F, double *F0, int nf)
{
    if (!c || !F || !F0 || nf < 1) return fail(c, -1, "st_filter: bad arguments");
    return chain3(c, st::K_FILT, F, F0, nf);
}

extern "C" int st_gauss(st_ctx *c, const double *F, double *Fg, int nf)
{
    if (!c || !F || !Fg || nf < 1) return fail(c, -1, "st_gauss: bad arguments");
    return chain3(c, st::K_GAUSS, F, Fg, nf);
}

extern "C" int st_rhs(st_ctx *c, const double *Q, double *Qrhs)
{
    if (!c) return -1;
    if (Q) TRY(up(c, c->Q, Q, 5));
    TRY(seq_rhs(c, base_args(c)));
    if (Qrhs) return down(c, Qrhs, c->dQ, 5);
    return 0;
}

extern "C" int st_hyper_rhs(st_ctx *c, double *Qrhs)
{
    if (!c) return -1;
    if (!c->p.hyper_param) return 0;                         // src/hyper_rhs.f:50-52
    if (Qrhs) TRY(up(c, c->dQ, Qrhs, 5));
    TRY(seq_hyper(c, base_args(c)));
    if (Qrhs) return down(c, Qrhs, c->dQ, 5);
    return 0;
}

extern "C" int st_compute_hyper(st_ctx *c, const double *Q)
{
    if (!c) return -1;
    if (Q) TRY(up(c, c->Q, Q, 5));
    if (c->p.hyper_param) TRY(seq_compute_hyper(c, 0, 1));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    return 0;
}

extern "C" int st_compute_grad(st_ctx *c, double *maxdil, double *enstrophy)
{
    if (!c) return -1;
    const long N = c->N, ni = c->ni;
    double *W = c->dQp;
    TRY(kerr(c, st::run_hyp_prep(c->Q, nullptr, N, ni, 0, W, c->s), "grad prep"));
    // /GRAD/ velocity gradients G[(3d + i) N] = d u_i / d x_d (src/grad.f:55-96)
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < 3; i++) TRY(op1(c, st::K_D1, d, W + i * N, c->G + (3 * d + i) * N, nullptr, 0, 1.0));
    TRY(kerr(c, st::run_gradred(c->G, N, ni, c->red, c->s), "grad reductions"));
    const int nb = st::gradred_blocks(ni);
    HIPCHK(c, hipMemcpyAsync(c->red_h, c->red, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    double m = c->red_h[0], t = c->red_h[1];
    for (int b = 1; b < nb; b++) {
        m = c->red_h[2 * b] > m ? c->red_h[2 * b] : m;
        t += c->red_h[2 * b + 1];
    }
    if (maxdil) *maxdil = m;                                               // Compute_Max (:104)
    if (enstrophy) *enstrophy = 0.5 * t / ((double)c->Mx * c->My * c->Mz);   // :113-115
    return 0;
}

extern "C" int st_get_hyper(st_ctx *c, double *H3)
{
    if (!c || !H3) return -1;
    return down(c, H3, c->Hc, 3);
}

extern "C" int st_set_hyper(st_ctx *c, const double *H3)
{
    if (!c || !H3) return -1;
    TRY(up(c, c->Hc, H3, 3));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int st_get_grad(st_ctx *c, double *G12)
{
    if (!c || !G12) return -1;
    return down(c, G12, c->G, 12);
}

extern "C" int st_fix_pressure(st_ctx *c, double *Q)
{
    if (!c || !Q) return -1;
    TRY(up(c, c->T5, Q, 5));
    st::PassArgs a = base_args(c);
    hipError_t err = st::run_fixp(a, c->T5, c->s);
    if (err != hipSuccess) return fail(c, -4, "fix_pressure: %s", hipGetErrorString(err));
    return down(c, Q, c->T5, 5);
}

extern "C" int st_cfl(st_ctx *c, double *stab, double *resv)
{
    if (!c) return -1;
    double v[2];
    TRY(reduce(c, 0, 2, v));
    if (stab) *stab = v[0];
    if (resv) *resv = v[1];
    return 0;
}

extern "C" int st_tstep(st_ctx *c, double cfl_resv, int step, double *dt)
{
    if (!c || !dt) return -1;
    double t = c->p.cfl_cal / cfl_resv;                      // src/lddrk4_periodic_sf.f:96
    // :99; the non-periodic driver drops the step > 1 guard
    // (src/lddrk4_nonperiodic_sf.f:105): before the first Hyper_RHS the
    // coefficients are zero and the limits infinite
    if (c->p.hyper_param == 1 && (step > 1 || c->g.nonperiodic_x)) {
        double v[3];
        TRY(reduce(c, 1, 3, v));
        const double p2 = (double)0.2f;                       // default-real literal 0.2 (:141)
        if (p2 * v[0] < t) t = p2 * v[0];
        if (p2 * v[1] < t) t = p2 * v[1];
        if (p2 * v[2] < t) t = p2 * v[2];
    }
    *dt = t;
    return 0;
}

extern "C" int st_rk_stage(st_ctx *c, double alpha, double beta, double dt, int reset_qw)
{
    if (!c) return -1;
    const st::PassArgs base = base_args(c);
    if (reset_qw) TRY(fill(c, c->QW, 5, base, 0.0));
    TRY(seq_stage(c, base, alpha, beta, dt));
    return 0;
}

extern "C" int st_step(st_ctx *c, double dt, double *time, int *step)
{
    if (!c) return -1;
    TRY(stage_loop(c, dt, time, step));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    return 0;
}

extern "C" int st_march(st_ctx *c, int nsteps, double *time, int *step, double *dt_out)
{
    if (!c || !time || !step) return -1;
    for (int it = 0; it < nsteps; it++) {
        double stab, resv, dt;
        TRY(st_cfl(c, &stab, &resv));
        TRY(st_tstep(c, resv, *step, &dt));
        if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, -5, "non-finite or non-positive dt %g at step %d", dt, *step);
        TRY(stage_loop(c, dt, time, step));
        if (dt_out) dt_out[it] = dt;
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->prof) prof_collect(c);
    return 0;
}

// ---------------------------------------------------------------------
// batches: same-size instances advanced together (no reference counterpart)
// ---------------------------------------------------------------------
namespace {

int slot_ok(st_ctx *c, int slot)
{
    if (slot < 0 || slot >= c->ns) return fail(c, -1, "slot %d out of range (batch of %d)", slot, c->ns);
    return 0;
}

// Upload the active flags and the list of active slots (the launches of a
// batch step cover the listed slots only).
int push_act(st_ctx *c)
{
    std::vector<int> list;
    for (int i = 0; i < c->ns; i++)
        if (c->act_h[i]) list.push_back(i);
    c->nact = (int)list.size();
    HIPCHK(c, hipMemcpy(c->s_act, c->act_h.data(), sizeof(int) * c->ns, hipMemcpyHostToDevice));
    if (c->nact) HIPCHK(c, hipMemcpy(c->s_list, list.data(), sizeof(int) * c->nact, hipMemcpyHostToDevice));
    return 0;
}

// Capture one batch step into a graph, once per active-slot count (the
// launch grids cover the active slots); the slot list and tables it reads
// live in device memory, so a graph stays valid as slots come and go.
int batch_graph(st_ctx *c)
{
    if (c->graphs[c->nact]) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed));
    const int r = batch_step(c);
    const hipError_t e = hipStreamEndCapture(c->s, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(c, -3, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(c, -3, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    c->graphs[c->nact] = x;
    return 0;
}

}  // namespace

extern "C" int st_slot_load(st_ctx *c, int slot, const double *Q, double Re, double time, int step, int nsteps)
{
    if (!c || !Q || nsteps < 0) return fail(c, -1, "st_slot_load: bad arguments");
    TRY(slot_ok(c, slot));
    TRY(up(c, c->Q, Q, 5, slot));
    // the slot's clock and target, 1/Re (src/define_periodic_sf.f:179-189),
    // and the artificial-transport coefficients of a start: Compute_hyper
    // before any RHS (src/main.f:100), so the dt limits of a slot resumed
    // at step > 1 (or of the non-periodic rule) see kappa* of its state
    const double re_inv = Re > 0.0 ? 1.0 / Re : 0.0;
    const int target = step + nsteps;
    HIPCHK(c, hipMemcpyAsync(c->s_re + slot, &re_inv, sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->s_time + slot, &time, sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->s_step + slot, &step, sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->s_target + slot, &target, sizeof(int), hipMemcpyHostToDevice, c->s));
    if (c->p.hyper_param) TRY(seq_compute_hyper(c, slot, 0));
    else
        for (int f = 0; f < 3; f++) {
            hipError_t e = st::run_fill(c->Hc + f * c->N + (long)slot * c->ni, c->ni, 0.0, c->s);
            if (e != hipSuccess) return fail(c, -4, "fill: %s", hipGetErrorString(e));
        }
    HIPCHK(c, hipStreamSynchronize(c->s));
    c->act_h[slot] = 1;
    return push_act(c);
}

extern "C" int st_slot_store(st_ctx *c, int slot, double *Q)
{
    if (!c || !Q) return fail(c, -1, "st_slot_store: bad arguments");
    TRY(slot_ok(c, slot));
    return down(c, Q, c->Q, 5, slot);
}

extern "C" int st_slot_release(st_ctx *c, int slot)
{
    if (!c) return -1;
    TRY(slot_ok(c, slot));
    HIPCHK(c, hipStreamSynchronize(c->s));
    c->act_h[slot] = 0;
    return push_act(c);
}

extern "C" int st_batch_march(st_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return -1;
    if (!c->nact) return 0;
    if (c->prof || !c->use_graph || getenv("ST_NO_GRAPH")) {
        for (int i = 0; i < nsteps; i++) TRY(batch_step(c));
        return 0;
    }
    TRY(batch_graph(c));
    for (int i = 0; i < nsteps; i++) HIPCHK(c, hipGraphLaunch(c->graphs[c->nact], c->s));
    return 0;
}

extern "C" int st_batch_status(st_ctx *c, int *step, double *time, double *dt_last, int *active)
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    const size_t ns = (size_t)c->ns;
    if (step) HIPCHK(c, hipMemcpy(step, c->s_step, sizeof(int) * ns, hipMemcpyDeviceToHost));
    if (time) HIPCHK(c, hipMemcpy(time, c->s_time, sizeof(double) * ns, hipMemcpyDeviceToHost));
    if (dt_last) HIPCHK(c, hipMemcpy(dt_last, c->s_dtlast, sizeof(double) * ns, hipMemcpyDeviceToHost));
    if (active) std::memcpy(active, c->act_h.data(), sizeof(int) * ns);
    return 0;
}

extern "C" int st_batch_slots(const st_ctx *c) { return c ? c->ns : 0; }

// ---------------------------------------------------------------------
// groups: batches of different grid sizes stepped together, one ragged
// launch per pass (no reference counterpart)
// ---------------------------------------------------------------------
struct st_group {
    st_ctx *m[4] = {nullptr, nullptr, nullptr, nullptr};   // by size-family member
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;                                 // orders the member streams with s
    std::map<std::array<int, 4>, hipGraphExec_t> graphs;     // one step per tuple of active-slot counts
};

namespace {

// Record one stage of member m (nothing is launched); the member's passes
// all take the M/2-point layouts of the size family.
int group_record(st_ctx *m, const st::PassArgs &a, double ra, double rb, std::vector<Rec> &out)
{
    out.clear();
    m->rec = &out;
    m->force_narrow = 1;
    const int r = seq_stage(m, a, ra, rb, 0.0);
    m->rec = nullptr;
    m->force_narrow = 0;
    return r;
}

// One step of every member's running slots, on the group's stream: per
// member the device dt (as batch_step), then the stages with each pass
// one ragged launch over all members.
int group_step(st_group *G)
{
    st::PassArgs a[4];
    int c0 = -1;
    hipError_t err = hipSuccess;
    for (int c = 0; c < 4; c++) {
        st_ctx *m = G->m[c];
        if (!m || !m->nact) continue;
        if (c0 < 0) c0 = c;
        st::PassArgs &b = a[c];
        b = base_args(m);
        b.nslot = m->nact;
        b.slots = m->s_list;
        b.act = m->s_act;
        int nb = 0;
        double *pc = m->red, *pt = m->red + (long)st::reduce_blocks(m->ni) * m->ns * 2;
        err = st::run_reduce(0, b, pc, &nb, G->s);
        if (err == hipSuccess && m->p.hyper_param == 1) err = st::run_reduce(1, b, pt, &nb, G->s);
        if (err == hipSuccess)
            err = st::run_dt(pc, pt, nb, m->ns, m->s_act, m->s_target, m->s_step, m->p.cfl_cal, m->p.hyper_param, 0,
                             m->s_run, m->s_dt, m->s_dtlast, G->s);
        if (err != hipSuccess) return st::set_call_error(-4, "group dt (%d points): %s", m->Mx, hipGetErrorString(err));
        b.act = m->s_run;
        b.sre = m->s_re;
        b.sdt = m->s_dt;
    }
    if (c0 < 0) return 0;
    const st_params &p = G->m[c0]->p;
    std::vector<Rec> rec[4];
    for (int loop = 0; loop < 2; loop++) {
        const double *ra, *rb;
        const int ns = rk_loop(p.rk_scheme, loop, &ra, &rb);
        if (!ns) break;
        for (int k = 0; k < ns; k++) {
            for (int c = c0; c < 4; c++) {
                st_ctx *m = G->m[c];
                if (!m || !m->nact) continue;
                if (group_record(m, a[c], ra[k], rb[k], rec[c]))
                    return st::set_call_error(-4, "group stage (%d points): %s", m->Mx, m->err);
                if (rec[c].size() != rec[c0].size())
                    return st::set_call_error(-4, "group stage: members record different pass sequences");
            }
            for (size_t i = 0; i < rec[c0].size(); i++) {
                st::RaggedSpec sp;
                std::memset(&sp, 0, sizeof(sp));
                const Rec &r0 = rec[c0][i];
                sp.dir = r0.dir;
                sp.first = c0;
                for (int c = c0; c < 4; c++) {
                    st_ctx *m = G->m[c];
                    if (!m || !m->nact) continue;
                    const Rec &r = rec[c][i];
                    if (r.kind != r0.kind || r.dir != r0.dir)
                        return st::set_call_error(-4, "group stage: members record different pass sequences");
                    sp.t[c] = r.t;
                    sp.My[c] = m->My;
                    sp.Mz[c] = m->Mz;
                    sp.Mx[c] = m->Mx;
                    sp.a[c] = r.a;
                }
                err = st::run_pass_ragged(r0.kind, sp, G->s);
                if (err != hipSuccess)
                    return st::set_call_error(-4, "group pass %s dir %d: %s", kind_names[r0.kind], r0.dir,
                                              hipGetErrorString(err));
            }
        }
        const int last = loop == 1 || !rk_loop(p.rk_scheme, 1, &ra, &rb);
        for (int c = c0; c < 4; c++) {
            st_ctx *m = G->m[c];
            if (!m || !m->nact) continue;
            err = st::run_clock(m->ns, m->s_run, m->s_dt, m->s_time, m->s_step, last, G->s);
            if (err != hipSuccess) return st::set_call_error(-4, "group clock: %s", hipGetErrorString(err));
        }
    }
    return 0;
}

#define GCHK(x)                                                                                    \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return st::set_call_error(-3, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Capture one group step per tuple of active-slot counts.
int group_graph(st_group *G, const std::array<int, 4> &key)
{
    if (G->graphs.count(key)) return 0;
    hipGraph_t g = nullptr;
    GCHK(hipStreamBeginCapture(G->s, hipStreamCaptureModeRelaxed));
    const int r = group_step(G);
    const hipError_t e = hipStreamEndCapture(G->s, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return st::set_call_error(-3, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return st::set_call_error(-3, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    G->graphs[key] = x;
    return 0;
}

}  // namespace

extern "C" int st_group_create(st_ctx *const *members, int nmembers, st_group **out)
{
    if (!members || !out || nmembers < 1 || nmembers > 4)
        return st::set_call_error(-1, "st_group_create: 1 to 4 member batches");
    *out = nullptr;
    st_group *G = new st_group();
    for (int i = 0; i < nmembers; i++) {
        st_ctx *m = members[i];
        if (!m) {
            delete G;
            return st::set_call_error(-1, "st_group_create: member %d is null", i);
        }
        const int c = st::ragged_member(m->Mx);
        if (c < 0 || m->My != m->Mx || m->Mz != m->Mx) {
            delete G;
            return st::set_call_error(-1, "st_group_create: a member grid must be 48^3, 64^3, 96^3 or 128^3 (got %dx%dx%d)",
                                      m->Mx, m->My, m->Mz);
        }
        if (G->m[c]) {
            delete G;
            return st::set_call_error(-1, "st_group_create: two members of %d^3 (one batch per size)", m->Mx);
        }
        if (m->g.nonperiodic_x || m->spg) {
            delete G;
            return st::set_call_error(-1, "st_group_create: members must be periodic, without a sponge");
        }
        const st_ctx *m0 = members[0];
        if (m->p.hyper_param != m0->p.hyper_param || m->p.rk_scheme != m0->p.rk_scheme) {
            delete G;
            return st::set_call_error(-1, "st_group_create: members must share hyper_param and rk_scheme");
        }
        for (int d = 0; d < 3; d++)
            if (m->th[3 + d].M != st::RAGGED_M[c] || m->th[3 + d].P != st::RAGGED_P[c]) {
                delete G;
                return st::set_call_error(-1, "st_group_create: %d^3 has no %d-point layout on %d lanes", m->Mx,
                                          st::RAGGED_M[c], st::RAGGED_P[c]);
            }
        G->m[c] = m;
    }
    hipError_t e = hipStreamCreateWithFlags(&G->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&G->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (G->s) (void)hipStreamDestroy(G->s);
        delete G;
        return st::set_call_error(-3, "st_group_create: %s", hipGetErrorString(e));
    }
    *out = G;
    return 0;
}

extern "C" int st_group_march(st_group *G, int nsteps)
{
    if (!G || nsteps < 0) return st::set_call_error(-1, "st_group_march: bad arguments");
    std::array<int, 4> key{0, 0, 0, 0};
    int any = 0;
    // the group's stream waits for the members' own work (loads, releases)
    for (int c = 0; c < 4; c++) {
        st_ctx *m = G->m[c];
        if (!m) continue;
        key[c] = m->nact;
        any |= m->nact;
        GCHK(hipEventRecord(G->ev, m->s));
        GCHK(hipStreamWaitEvent(G->s, G->ev, 0));
    }
    if (!any || !nsteps) return 0;
    if (getenv("ST_NO_GRAPH")) {
        for (int i = 0; i < nsteps; i++) TRY(group_step(G));
    } else {
        TRY(group_graph(G, key));
        const hipGraphExec_t x = G->graphs[key];
        for (int i = 0; i < nsteps; i++) GCHK(hipGraphLaunch(x, G->s));
    }
    // and the members' later work waits for the steps
    GCHK(hipEventRecord(G->ev, G->s));
    for (int c = 0; c < 4; c++)
        if (G->m[c]) GCHK(hipStreamWaitEvent(G->m[c]->s, G->ev, 0));
    return 0;
}

extern "C" int st_group_sync(st_group *G)
{
    if (!G) return st::set_call_error(-1, "st_group_sync: null group");
    GCHK(hipStreamSynchronize(G->s));
    return 0;
}

extern "C" void st_group_destroy(st_group *G)
{
    if (!G) return;
    (void)hipStreamSynchronize(G->s);
    for (auto &kv : G->graphs) (void)hipGraphExecDestroy(kv.second);
    if (G->ev) (void)hipEventDestroy(G->ev);
    if (G->s) (void)hipStreamDestroy(G->s);
    delete G;
}

extern "C" int st_query(st_ctx *c)
{
    if (!c) return -1;
    const hipError_t e = hipStreamQuery(c->s);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(c, -3, "hipStreamQuery: %s", hipGetErrorString(e));
}

// Pick each pass's lane layout (M or M/2 points per lane) and tile width
// (256, 512, 1024 threads) by timing: every candidate in turn runs `reps`
// RK stages over all slots with per-pass HIP events, and each (pass,
// direction) keeps its fastest valid candidate.  The state and the
// artificial-transport coefficients are restored afterwards.
extern "C" int st_autotune(st_ctx *c, int reps)
{
    if (!c || reps < 1) return fail(c, -1, "st_autotune: bad arguments");
    HIPCHK(c, hipStreamSynchronize(c->s));
    prof_collect(c);
    double *bak = nullptr;
    HIPCHK(c, hipMalloc(&bak, sizeof(double) * 8 * c->N));
    const size_t f5 = sizeof(double) * 5 * c->N, f3 = sizeof(double) * 3 * c->N;
    hipError_t e = hipMemcpyAsync(bak, c->Q, f5, hipMemcpyDeviceToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(bak + 5 * c->N, c->Hc, f3, hipMemcpyDeviceToDevice, c->s);
    if (e != hipSuccess) {
        (void)hipFree(bak);
        return fail(c, -3, "autotune backup: %s", hipGetErrorString(e));
    }
    int dn[NKIND][3], dt[NKIND][3], bn[NKIND][3], bt[NKIND][3];
    double best[NKIND][3];
    std::memcpy(dn, c->cfg_narrow, sizeof(dn));
    std::memcpy(dt, c->cfg_nt, sizeof(dt));
    std::memcpy(bn, dn, sizeof(bn));
    std::memcpy(bt, dt, sizeof(bt));
    for (int k = 0; k < NKIND; k++)
        for (int d = 0; d < 3; d++) best[k][d] = INFINITY;
    static const int cand[6][2] = {{0, 256}, {0, 512}, {0, 1024}, {1, 256}, {1, 512}, {1, 1024}};
    st::PassArgs base = base_args(c);
    base.nslot = c->ns;
    st::PassArgs dry = base;
    dry.dry = 1;
    const double *ra, *rb;
    rk_loop(c->p.rk_scheme, 0, &ra, &rb);
    const int prof_was = c->prof;
    int rc = 0;
    for (int ci = 0; ci < 6 && !rc; ci++) {
        int valid[NKIND][3] = {};
        for (int k = st::K_GRAD; k < st::K_NPASS; k++)
            for (int d = 0; d < 3; d++) {
                c->cfg_narrow[k][d] = cand[ci][0];
                c->cfg_nt[k][d] = cand[ci][1];
                valid[k][d] = pass(c, k, d, dry) == 0;
                if (!valid[k][d]) {
                    c->cfg_narrow[k][d] = dn[k][d];
                    c->cfg_nt[k][d] = dt[k][d];
                }
            }
        c->err[0] = 0;
        c->prof = 1;
        std::memset(c->prof_ms, 0, sizeof(c->prof_ms));
        std::memset(c->prof_cnt, 0, sizeof(c->prof_cnt));
        for (int r = 0; r < reps && !rc; r++) rc = seq_stage(c, base, ra[0], rb[0], 1.0e-6);
        if (hipStreamSynchronize(c->s) != hipSuccess && !rc) rc = fail(c, -3, "autotune: stream error");
        prof_collect(c);
        for (int k = st::K_GRAD; k < st::K_NPASS; k++)
            for (int d = 0; d < 3; d++) {
                const int pk = k * 3 + d;
                if (!valid[k][d] || !c->prof_cnt[pk]) continue;
                const double ms = c->prof_ms[pk] / (double)c->prof_cnt[pk];
                if (ms < best[k][d]) {
                    best[k][d] = ms;
                    bn[k][d] = cand[ci][0];
                    bt[k][d] = cand[ci][1];
                }
            }
    }
    std::memcpy(c->cfg_narrow, rc ? dn : bn, sizeof(bn));
    std::memcpy(c->cfg_nt, rc ? dt : bt, sizeof(bt));
    c->prof = prof_was;
    std::memset(c->prof_ms, 0, sizeof(c->prof_ms));
    std::memset(c->prof_cnt, 0, sizeof(c->prof_cnt));
    e = hipMemcpyAsync(c->Q, bak, f5, hipMemcpyDeviceToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->Hc, bak + 5 * c->N, f3, hipMemcpyDeviceToDevice, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    (void)hipFree(bak);
    for (auto &g : c->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, -3, "autotune restore: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int st_sponge(st_ctx *c, double amp, double power, double frac, const double ref_left[5],
                         const double ref_right[5])
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (amp == 0.0) {
        if (c->spg) (void)hipFree(c->spg);
        c->spg = nullptr;
    } else {
        if (!ref_left || !ref_right || !(frac > 0.0) || frac >= 1.0) return fail(c, -1, "st_sponge: bad arguments");
        std::vector<double> h(c->Mx + 10);
        for (int i = 0; i < c->Mx; i++) h[i] = st::sponge_sigma(c->Mx, i, amp, power, frac);
        for (int l = 0; l < 5; l++) {
            h[c->Mx + l] = ref_left[l];
            h[c->Mx + 5 + l] = ref_right[l];
        }
        c->spg_left = (int)std::ceil((frac * c->Mx) / 2.0);
        if (!c->spg) HIPCHK(c, hipMalloc(&c->spg, sizeof(double) * (c->Mx + 10)));
        HIPCHK(c, hipMemcpy(c->spg, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    }
    for (auto &g : c->graphs)                                  // the captured steps change
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return 0;
}

extern "C" int st_ic_shock(int nx, int ny, int nz, const double period[3], double mach, double frac, double width,
                           double k0, double urms, double gamma, unsigned long long seed, double *Q, double *states)
{
    if (!Q || !period) return -1;
    return st::ic_shock(nx, ny, nz, period, mach, frac, width, k0, urms, gamma, seed, Q, states);
}

extern "C" int st_pass_config(st_ctx *c, int kind_dir, int *narrow, int *nt)
{
    if (!c || kind_dir < 0 || kind_dir >= NKIND * 3) return -1;
    if (narrow) *narrow = c->cfg_narrow[kind_dir / 3][kind_dir % 3];
    if (nt) *nt = c->cfg_nt[kind_dir / 3][kind_dir % 3];
    return 0;
}

extern "C" int st_pass_config_set(st_ctx *c, int kind_dir, int narrow, int nt)
{
    if (!c || kind_dir < 0 || kind_dir >= NKIND * 3 || (narrow != 0 && narrow != 1) ||
        (nt != 256 && nt != 512 && nt != 1024))
        return -1;
    c->cfg_narrow[kind_dir / 3][kind_dir % 3] = narrow;
    c->cfg_nt[kind_dir / 3][kind_dir % 3] = nt;
    for (auto &g : c->graphs)                                  // captured with the old shapes
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return 0;
}

extern "C" int st_diagnostics(st_ctx *c, double *out7)
{
    if (!c || !out7) return -1;
    double v[7];
    TRY(reduce(c, 2, 7, v));
    const double NN = (double)(c->Mx * c->My * c->Mz);
    out7[0] = v[0] / NN;
    out7[1] = v[1] / NN;
    out7[2] = v[2] / NN;
    out7[3] = 0.5 * v[3] / NN;
    out7[4] = v[4] / NN;
    out7[5] = v[5] / NN;
    out7[6] = v[6] / NN;
    return 0;
}

extern "C" int st_profile(st_ctx *c, int on)
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    prof_collect(c);
    c->prof = on ? 1 : 0;
    std::memset(c->prof_ms, 0, sizeof(c->prof_ms));
    std::memset(c->prof_cnt, 0, sizeof(c->prof_cnt));
    return 0;
}

extern "C" int st_profile_read(st_ctx *c, int nmax, int *count, double *total_ms, double *fields_rw)
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    prof_collect(c);
    const int n = nmax < NKIND * 3 ? nmax : NKIND * 3;
    for (int i = 0; i < n; i++) {
        if (count) count[i] = (int)c->prof_cnt[i];
        if (total_ms) total_ms[i] = c->prof_ms[i];
        if (fields_rw) fields_rw[i] = c->prof_fields[i];
    }
    return n;
}

extern "C" const char *st_pass_name(int kind)
{
    static char buf[NKIND * 3][32];
    if (kind < 0 || kind >= NKIND * 3) return "";
    if (!buf[kind][0]) snprintf(buf[kind], sizeof(buf[kind]), "%s.%c", kind_names[kind / 3], "xyz"[kind % 3]);
    return buf[kind];
}

extern "C" int st_ic_isotropic(int nx, int ny, int nz, const double period[3], double k0, double urms,
                               double mt, double gamma, unsigned long long seed, double *Q)
{
    if (!Q || !period) return -1;
    return st::ic_isotropic(nx, ny, nz, period, k0, urms, mt, gamma, seed, Q);
}
