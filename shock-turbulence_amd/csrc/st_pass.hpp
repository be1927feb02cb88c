// Line-pass kernel template and its launchers, shared by the pass
// translation units (st_pass_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "st_engine.hpp"
#include "st_kernels.h"

namespace st {

// ---------------------------------------------------------------------
// generic pass kernel
// ---------------------------------------------------------------------
template <class PASS>
constexpr int pass_slots() { return PASS::REUSE ? PASS::NSLOT : PASS::NSLOT + PASS::NEARLY; }

// NCARRY of a pass (0 when it names none): values per point that
// store_early() hands to store() in registers (REUSE passes only).
template <class T, class = void>
struct carry_n { static constexpr int value = 0; };
template <class T>
struct carry_n<T, std::void_t<decltype(T::NCARRY)>> { static constexpr int value = T::NCARRY; };

// MID of a pass (false when it names none): a second compute phase.  After
// compute(), mid() rewrites the NSLOT staged values of each point in memory
// order (it may read and write global fields of its own points; it also gets
// the point's index along the line), then compute2() runs on the rewritten
// slots in the line mapping.
// NKEEP of a pass (0 when it names none): the last NKEEP staged input
// slots, which compute() only reads, are handed to store() after the late
// results, so the store phase need not read those fields again.
template <class T, class = void>
struct keep_n { static constexpr int value = 0; };
template <class T>
struct keep_n<T, std::void_t<decltype(T::NKEEP)>> { static constexpr int value = T::NKEEP; };

// NSF / NSFE of a pass (0 when it names none): global values per point
// that its late (store) or early (store_early) store phase reads.  pass_body
// loads them with sfetch() / sfetch_early() for SFB points (SF_BATCH, default
// all M) before the first store of the batch, so that the loads are in
// flight together instead of each waiting behind the stores before it (the
// compiler may not move a load above a store to a pointer that could alias
// it).  store() / store_early() then take them as a last argument.
template <class T, class = void>
struct sf_n { static constexpr int value = 0; static constexpr int batch = 0; };
template <class T>
struct sf_n<T, std::void_t<decltype(T::NSF)>> { static constexpr int value = T::NSF; static constexpr int batch = T::SF_BATCH; };
template <class T, class = void>
struct sfe_n { static constexpr int value = 0; };
template <class T>
struct sfe_n<T, std::void_t<decltype(T::NSFE)>> { static constexpr int value = T::NSFE; };

// lds_add(M) of a pass (false when it names none): the additive LDS layout
// (lds_idx) instead of the swizzled one, for the pass at M points per lane.
template <class T, int M, class = void>
struct lds_add { static constexpr bool value = false; };
template <class T, int M>
struct lds_add<T, M, std::void_t<decltype(T::lds_add(M))>> { static constexpr bool value = T::lds_add(M); };

// NHOLD of a pass (0 when it names none): values per point that load()
// returns after the NSLOT staged ones and that wait in registers while
// compute_a() runs on the slots; they are then written into the slots
// hold_slot(j) (which compute_a() has finished reading) and compute() runs.
// A pass can so stage more fields over its compute than it has slots.
template <class T, class = void>
struct hold_n { static constexpr int value = 0; };
template <class T>
struct hold_n<T, std::void_t<decltype(T::NHOLD)>> { static constexpr int value = T::NHOLD; };

// NLCARRY of a pass (0 when it names none): values per point that load()
// returns after the staged and held ones and that wait in registers until
// the memory-order phase that uses them: mid() for a pass with MID, else
// the late store (either takes them as its last argument).  These phases
// visit the same points per thread as the load, so a value the load phase
// reads or derives at a point need not be read again there.
template <class T, class = void>
struct lcarry_n { static constexpr int value = 0; };
template <class T>
struct lcarry_n<T, std::void_t<decltype(T::NLCARRY)>> { static constexpr int value = T::NLCARRY; };

template <class T, class = void>
struct has_mid { static constexpr bool value = false; };
template <class T>
struct has_mid<T, std::void_t<decltype(T::MID)>> { static constexpr bool value = T::MID; };

// Slots 0..NSLOT-1 hold the staged inputs; compute() may write final
// results early (freeing their registers) into slots NSLOT..NSLOT+NEARLY-1,
// or, with REUSE, over its own points of the input slots 0..NEARLY-1 once
// it no longer reads them (each lane reads and writes only its own points
// of the tile during compute).  The NOUT register results are staged into
// slots 0..NOUT-1 after compute.  Without REUSE, store() receives per point
// the NOUT late results followed by the NEARLY early ones; with REUSE the
// early ones go first through store_early() and store() gets the late ones.
// A pass may also name NPF values per point that its store phase reads
// (read-modify-write operands): prefetch() loads them during the load phase,
// which visits the same points per thread as the store phase, and they
// reach store_early() (REUSE) or store() appended after the results.
// A REUSE pass with NCARRY > 0 gets a per-point register array in both
// store phases (same points per thread): store_early() fills it, store()
// reads it, so early results can feed the late store without a slot.
template <int M, int P, int NT, bool NP, class PASS>
DEV void pass_body(const Geo &g, const PassArgs &a0, long tile)
{
    extern __shared__ double lds[];
    constexpr int NS = PASS::NSLOT, NO = PASS::NOUT, NE = PASS::NEARLY;
    constexpr bool RE = PASS::REUSE;
    static_assert(NO <= NS, "late outputs must fit in the input slots");
    static_assert(NO > 0 || RE, "a pass without late outputs stores through store_early");
    static_assert(!RE || NE <= NS, "early outputs must fit in the input slots");
    constexpr int NPF = PASS::NPF;
    constexpr bool AL = lds_add<PASS, M>::value;
    constexpr int NH = hold_n<PASS>::value, NLC = lcarry_n<PASS>::value;
    static_assert(NPF == 0 || PASS::LOAD_UNROLL >= M, "prefetching needs the load loop unrolled");
    static_assert(NH + NLC == 0 || PASS::LOAD_UNROLL >= M, "holding needs the load loop unrolled");
    double pf[M][NPF > 0 ? NPF : 1];
    double hold[M][NH > 0 ? NH : 1];
    double lc[M][NLC > 0 ? NLC : 1];
    constexpr int NCR = carry_n<PASS>::value;
    static_assert(NCR == 0 || (RE && NO > 0), "a carry goes from store_early to store");
    double cr[M][NCR > 0 ? NCR : 1];
    const int si = (int)(tile / g.tpi);
    const int slot = a0.slots ? a0.slots[si] : si;
    if (a0.act && !a0.act[slot]) return;          // uniform per workgroup, before any barrier
    PassArgs a;
    slot_params(a0, slot, a);
    const Walk<M, P, NT> W(g, tile - (long)si * g.tpi, (long)slot * a.ni, threadIdx.x);
    constexpr int ss = NT * M;           // slot size: B*M*P with B*P = NT

    // 1. load (memory order, coalesced)
#pragma unroll PASS::LOAD_UNROLL
    for (int r = 0; r < M; r++) {
        int b, s;
        const long gi = W.at(g, r, b, s);
        double v[NS + NH + NLC > 0 ? NS + NH + NLC : 1];
        PASS::load(g, a, gi, v);
        if constexpr (NPF > 0) PASS::prefetch(g, a, gi, pf[r]);
        const int k = s % M, c = s / M;
        const int li = lds_idx<M, AL>(b, k, c, P, lds_g<M>(b, NT / P));
#pragma unroll
        for (int q = 0; q < NS; q++) lds[q * ss + li] = v[q];
#pragma unroll
        for (int j = 0; j < NH; j++) hold[r][j] = v[NS + j];
#pragma unroll
        for (int j = 0; j < NLC; j++) lc[r][j] = v[NS + NH + j];
    }
    __syncthreads();

    // 2. compute (line mapping)
    const Lane<M, P, NP, AL> L = make_lane<M, P, NP, AL>(lds, ss);
    double out[NO > 0 ? NO : 1][M];
    if constexpr (NH > 0) {
        PASS::template compute_a<M, P, NP>(g, a, L, out);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < M; r++) {
            int b, s;
            (void)W.at(g, r, b, s);
            const int li = lds_idx<M, AL>(b, s % M, s / M, P, lds_g<M>(b, NT / P));
#pragma unroll
            for (int j = 0; j < NH; j++) lds[PASS::hold_slot(j) * ss + li] = hold[r][j];
        }
        __syncthreads();
    }
    PASS::template compute<M, P, NP>(g, a, L, out);
    // The memory-order walk again for the phases after the compute phase,
    // from a copy of the thread index that the compiler cannot see through:
    // the load phase's addresses and LDS indices are then computed anew
    // instead of being kept (in registers, or spilled) through the compute
    // phase.
    int tl = threadIdx.x;
    asm volatile("" : "+v"(tl));
    const Walk<M, P, NT> Wl(g, tile - (long)si * g.tpi, (long)slot * a.ni, tl);
    if constexpr (has_mid<PASS>::value) {
        static_assert(RE && NO == 0, "a second compute phase works on the slots");
        __syncthreads();
#pragma unroll
        for (int r = 0; r < M; r++) {
            int b, s;
            const long gi = Wl.at(g, r, b, s);
            const int li = lds_idx<M, AL>(b, s % M, s / M, P, lds_g<M>(b, NT / P));
            double v[NS];
#pragma unroll
            for (int q = 0; q < NS; q++) v[q] = lds[q * ss + li];
            if constexpr (NLC > 0) PASS::mid(g, a, gi, s, v, lc[r]);
            else PASS::mid(g, a, gi, s, v);
#pragma unroll
            for (int q = 0; q < NS; q++) lds[q * ss + li] = v[q];
        }
        __syncthreads();
        PASS::template compute2<M, P, NP>(g, a, L);
    }

    // 3. store (through LDS, memory order)
    constexpr int NSFE = sfe_n<PASS>::value, NSF = sf_n<PASS>::value;
    static_assert((NSF == 0 && NSFE == 0) || (RE && NCR > 0 && NPF == 0),
                  "store-phase fetches: passes with a carry, without prefetch");
    __syncthreads();
    if (RE) {
        double sfe[M][NSFE > 0 ? NSFE : 1];
        if constexpr (NSFE > 0) {
#pragma unroll
            for (int r = 0; r < M; r++) {
                int b, s;
                PASS::sfetch_early(g, a, Wl.at(g, r, b, s), sfe[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < M; r++) {
            int b, s;
            const long gi = Wl.at(g, r, b, s);
            const int li = lds_idx<M, AL>(b, s % M, s / M, P, lds_g<M>(b, NT / P));
            double v[NE + NPF > 0 ? NE + NPF : 1];
#pragma unroll
            for (int j = 0; j < NE; j++) v[j] = lds[j * ss + li];
#pragma unroll
            for (int j = 0; j < NPF; j++) v[NE + j] = pf[r][j];
            if constexpr (NSFE > 0) PASS::store_early(g, a, gi, v, cr[r], sfe[r]);
            else if constexpr (NCR > 0) PASS::store_early(g, a, gi, v, cr[r]);
            else PASS::store_early(g, a, gi, v);
        }
        if constexpr (NO == 0) return;
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NO; j++) {
#pragma unroll
        for (int k = 0; k < M; k++) L.st(j, k, out[j][k]);
    }
    __syncthreads();
    constexpr int SFB = NSF > 0 && sf_n<PASS>::batch < M ? sf_n<PASS>::batch : M;
    static_assert(SFB > 0, "store-fetch batch");
#pragma unroll
    for (int r0 = 0; r0 < M; r0 += SFB) {
        double sf[SFB][NSF > 0 ? NSF : 1];
        if constexpr (NSF > 0) {
#pragma unroll
            for (int j = 0; j < SFB; j++) {
                if (r0 + j >= M) break;
                int b, s;
                PASS::sfetch(g, a, Wl.at(g, r0 + j, b, s), sf[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < SFB; j++) {
            if (r0 + j >= M) break;
            const int r = r0 + j;
            int b, s;
            const long gi = Wl.at(g, r, b, s);
            const int k = s % M, c = s / M;
            const int li = lds_idx<M, AL>(b, k, c, P, lds_g<M>(b, NT / P));
            constexpr int NV = RE ? NO : NO + NE;
            constexpr int NK = keep_n<PASS>::value;
            static_assert(NK == 0 || (RE && NS - NK >= NO && NS - NK >= NE), "kept slots hold no results");
            double v[NV + NK + NPF > 0 ? NV + NK + NPF : 1];
#pragma unroll
            for (int q = 0; q < NO; q++) v[q] = lds[q * ss + li];
            if (!RE) {
#pragma unroll
                for (int q = NO; q < NV; q++) v[q] = lds[(NS + q - NO) * ss + li];
            }
#pragma unroll
            for (int q = 0; q < NK; q++) v[NV + q] = lds[(NS - NK + q) * ss + li];
#pragma unroll
            for (int q = 0; q < NPF; q++) v[NV + NK + q] = pf[r][q];
            constexpr bool LCS = NLC > 0 && !has_mid<PASS>::value;
            static_assert(!LCS || (NSF > 0 && NCR > 0), "a load carry to the store goes with store-phase fetches and a carry");
            if constexpr (LCS) PASS::store(g, a, gi, v, cr[r], sf[j], lc[r]);
            else if constexpr (NSF > 0) PASS::store(g, a, gi, v, cr[r], sf[j]);
            else if constexpr (NCR > 0) PASS::store(g, a, gi, v, cr[r]);
            else PASS::store(g, a, gi, v);
        }
    }
}

template <int M, int P, int NT, bool NP, class PASS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PASS::MIN_WAVES)))
k_pass(Geo g, PassArgs a0)
{
    pass_body<M, P, NT, NP, PASS>(g, a0, tile_of(g, blockIdx.x));
}

// Passes whose early outputs have their own slots.
struct NoReuse {
    static constexpr bool REUSE = false;
    static constexpr int MIN_WAVES = 1;
    static constexpr int LOAD_UNROLL = 8;
    static constexpr int NPF = 0;
    DEV static void prefetch(const Geo &, const PassArgs &, long, double *) {}
    DEV static void store_early(const Geo &, const PassArgs &, long, const double *) {}
};

// ---------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------
DEV double qf(const PassArgs &a, int comp, long gi) { return a.Q[comp * a.N + gi]; }

template <int M, int P, bool NP, bool AL>
DEV void qown(const Lane<M, P, NP, AL> &L, int k, double &q1, double &q2, double &q3, double &q4, double &q5)
{
    q1 = L.ld(0, k);
    q2 = L.ld(1, k);
    q3 = L.ld(2, k);
    q4 = L.ld(3, k);
    q5 = L.ld(4, k);
}

DEV void upd(double *p, double v, bool first)
{
    *p = first ? v : *p + v;
}


// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
template <int M, int P, int NT, bool NP, class PASS>
static hipError_t launch_t(const Geo &g, const PassArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)pass_slots<PASS>() * NT * M * sizeof(double);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (a.dry) return hipSuccess;
    static size_t lds_set = 0;
    if (lds > 64 * 1024 && lds > lds_set) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pass<M, P, NT, NP, PASS>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    hipLaunchKernelGGL((k_pass<M, P, NT, NP, PASS>), dim3((unsigned)g.ntiles), dim3(NT), lds, s, g, a);
    return hipGetLastError();
}

// The lane layouts (M points per lane, P lanes per line) that
// choose_layout and its narrow variant (M/2, 2P) produce for lines of
// 16..512 points (powers of two), 48..384 (three times a power of two),
// 40..320 (five times one) and 56..448 (seven times one).
template <class PASS, int NT, bool NP>
static hipError_t launch_m(int M, int P, const Geo &g, const PassArgs &a, hipStream_t s)
{
    if constexpr (NT == 1024) {
        // 16 or 32 lines per tile: lanes of 32 or 64 per line only
        switch (M * 100 + P) {
        case 432: return launch_t<4, 32, NT, NP, PASS>(g, a, s);
        case 464: return launch_t<4, 64, NT, NP, PASS>(g, a, s);
        case 832: return launch_t<8, 32, NT, NP, PASS>(g, a, s);
        case 864: return launch_t<8, 64, NT, NP, PASS>(g, a, s);
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (M * 100 + P) {
        case 116: return launch_t<1, 16, NT, NP, PASS>(g, a, s);
        case 208: return launch_t<2, 8, NT, NP, PASS>(g, a, s);
        case 216: return launch_t<2, 16, NT, NP, PASS>(g, a, s);
        case 408: return launch_t<4, 8, NT, NP, PASS>(g, a, s);
        case 416: return launch_t<4, 16, NT, NP, PASS>(g, a, s);
        case 432: return launch_t<4, 32, NT, NP, PASS>(g, a, s);
        case 464: return launch_t<4, 64, NT, NP, PASS>(g, a, s);
        case 808: return launch_t<8, 8, NT, NP, PASS>(g, a, s);
        case 816: return launch_t<8, 16, NT, NP, PASS>(g, a, s);
        case 832: return launch_t<8, 32, NT, NP, PASS>(g, a, s);
        case 864: return launch_t<8, 64, NT, NP, PASS>(g, a, s);
        // lines of 40, 80, 160, 320 (M = 5) and 56, 112, 224, 448 (M = 7) points
        case 508: return launch_t<5, 8, NT, NP, PASS>(g, a, s);
        case 516: return launch_t<5, 16, NT, NP, PASS>(g, a, s);
        case 532: return launch_t<5, 32, NT, NP, PASS>(g, a, s);
        case 564: return launch_t<5, 64, NT, NP, PASS>(g, a, s);
        case 708: return launch_t<7, 8, NT, NP, PASS>(g, a, s);
        case 716: return launch_t<7, 16, NT, NP, PASS>(g, a, s);
        case 732: return launch_t<7, 32, NT, NP, PASS>(g, a, s);
        case 764: return launch_t<7, 64, NT, NP, PASS>(g, a, s);
        // lines of 48, 96, 192 and 384 points
        case 316: return launch_t<3, 16, NT, NP, PASS>(g, a, s);
        case 332: return launch_t<3, 32, NT, NP, PASS>(g, a, s);
        case 364: return launch_t<3, 64, NT, NP, PASS>(g, a, s);
        case 608: return launch_t<6, 8, NT, NP, PASS>(g, a, s);
        case 616: return launch_t<6, 16, NT, NP, PASS>(g, a, s);
        case 632: return launch_t<6, 32, NT, NP, PASS>(g, a, s);
        case 664: return launch_t<6, 64, NT, NP, PASS>(g, a, s);
        default: return hipErrorInvalidValue;
        }
    }
}

// t holds the tables of the chosen layout (M points per lane, P lanes per
// line); nt threads per workgroup give B = nt/P lines per tile.
// XL = false for passes that never sweep x (no non-periodic instances).
// The tile geometry of a launch; false when nt threads cannot tile the lines.
static inline bool pass_geo(const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, Geo &g)
{
    g.dir = dir;
    g.n = t.n;
    g.My = My;
    g.Mz = Mz;
    g.Mx = Mx;
    g.P = t.P;
    g.B = nt / t.P;
    if (g.B < 1 || g.B * t.P != nt) return false;
    const long lot = (long)My * Mz * Mx / g.n;          // lines per instance
    if (lot % g.B || a.ni != (long)My * Mz * Mx || a.nslot < 1) return false;
    g.tpi = lot / g.B;
    g.ntiles = g.tpi * a.nslot;
    return true;
}

// One periodic layout and tile width of a pass (M, P, NT fixed).
template <class PASS, int M, int P, int NT>
static hipError_t launch_at(const DirTab &t, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    Geo g;
    if (t.M != M || t.P != P || t.nonper || !pass_geo(t, NT, dir, My, Mz, Mx, a, g)) return hipErrorInvalidValue;
    return launch_t<M, P, NT, false, PASS>(g, a, s);
}

template <class PASS, bool XL = true>
static hipError_t launch(const DirTab &t, int nt, int dir, int My, int Mz, int Mx, const PassArgs &a, hipStream_t s)
{
    Geo g;
    if (!pass_geo(t, nt, dir, My, Mz, Mx, a, g)) return hipErrorInvalidValue;
    // non-periodic lines are the x lines of the shock-tube link set
    if (t.nonper && dir == 0) {
        if constexpr (XL) {
            if (nt == 256) return launch_m<PASS, 256, true>(t.M, t.P, g, a, s);
            if (nt == 512) return launch_m<PASS, 512, true>(t.M, t.P, g, a, s);
            if (nt == 1024) return launch_m<PASS, 1024, true>(t.M, t.P, g, a, s);
        }
        return hipErrorInvalidValue;
    }
    if (nt == 256) return launch_m<PASS, 256, false>(t.M, t.P, g, a, s);
    if (nt == 512) return launch_m<PASS, 512, false>(t.M, t.P, g, a, s);
    if (nt == 1024) return launch_m<PASS, 1024, false>(t.M, t.P, g, a, s);
    return hipErrorInvalidValue;
}


// ---------------------------------------------------------------------
// Ragged launches: one launch of a pass over the instances of several
// grid sizes, each size a batch of its own (a group).  The sizes belong to
// a family fixed at compile time, each with its own lane layout; class c
// (the family member) owns workgroups [start[c], start[c+1]), padded to a
// multiple of 8 so that a class's consecutive tiles keep to one XCD.
// ---------------------------------------------------------------------
constexpr int RG_MAX = 4;
template <int M_, int P_>
struct Lay {
    static constexpr int M = M_, P = P_;
};
// The ensemble and serving family (RAGGED_N, RAGGED_M, RAGGED_P).
struct FamilyServe {
    using L0 = Lay<RAGGED_M[0], RAGGED_P[0]>;
    using L1 = Lay<RAGGED_M[1], RAGGED_P[1]>;
    using L2 = Lay<RAGGED_M[2], RAGGED_P[2]>;
    using L3 = Lay<RAGGED_M[3], RAGGED_P[3]>;
    static int M_of(int c) { return RAGGED_M[c]; }
    static int P_of(int c) { return RAGGED_P[c]; }
};

struct RaggedArgs {
    Geo g[RG_MAX];
    PassArgs a[RG_MAX];
    long start[RG_MAX + 1];
};
static_assert(sizeof(RaggedArgs) <= 4000, "kernel arguments");

template <int NT, class PASS, class F>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PASS::MIN_WAVES)))
k_ragged(RaggedArgs r)
{
    const long wg = blockIdx.x;
    int c = 0;
#pragma unroll
    for (int i = 1; i < RG_MAX; i++) c += wg >= r.start[i] ? 1 : 0;
    const long local = wg - r.start[c];
    switch (c) {
    case 0:
        if (local < r.g[0].ntiles)
            pass_body<F::L0::M, F::L0::P, NT, false, PASS>(r.g[0], r.a[0], tile_of(r.g[0], local));
        break;
    case 1:
        if (local < r.g[1].ntiles)
            pass_body<F::L1::M, F::L1::P, NT, false, PASS>(r.g[1], r.a[1], tile_of(r.g[1], local));
        break;
    case 2:
        if (local < r.g[2].ntiles)
            pass_body<F::L2::M, F::L2::P, NT, false, PASS>(r.g[2], r.a[2], tile_of(r.g[2], local));
        break;
    default:
        if (local < r.g[3].ntiles)
            pass_body<F::L3::M, F::L3::P, NT, false, PASS>(r.g[3], r.a[3], tile_of(r.g[3], local));
        break;
    }
}

// One ragged launch of PASS: per family member c, the tables t[c] (null:
// no instances of that size), the grid and the pass arguments.  Every
// member present must have its family layout and the periodic tables.
template <class PASS, class F = FamilyServe>
static hipError_t launch_ragged(const RaggedSpec &sp, hipStream_t s)
{
    constexpr int NT = 256;
    RaggedArgs r;
    long wg = 0;
    size_t lds = 0;
    for (int c = 0; c < RG_MAX; c++) {
        r.start[c] = wg;
        Geo &g = r.g[c];
        g = Geo();
        r.a[c] = sp.a[c];
        g.ntiles = 0;
        if (!sp.t[c] || sp.a[c].nslot < 1) continue;
        const DirTab &t = *sp.t[c];
        if (t.M != F::M_of(c) || t.P != F::P_of(c) || t.nonper) return hipErrorInvalidValue;
        g.dir = sp.dir;
        g.n = t.n;
        g.My = sp.My[c];
        g.Mz = sp.Mz[c];
        g.Mx = sp.Mx[c];
        g.P = t.P;
        g.B = NT / t.P;
        const long lot = (long)g.My * g.Mz * g.Mx / g.n;
        if (lot % g.B || sp.a[c].ni != (long)g.My * g.Mz * g.Mx) return hipErrorInvalidValue;
        g.tpi = lot / g.B;
        g.ntiles = g.tpi * sp.a[c].nslot;
        wg += (g.ntiles + 7) / 8 * 8;
        const size_t l = (size_t)pass_slots<PASS>() * NT * t.M * sizeof(double);
        lds = l > lds ? l : lds;
    }
    r.start[RG_MAX] = wg;
    if (wg == 0) return hipSuccess;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    static size_t lds_set = 0;
    if (lds > 64 * 1024 && lds > lds_set) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ragged<NT, PASS, F>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    hipLaunchKernelGGL((k_ragged<NT, PASS, F>), dim3((unsigned)wg), dim3(NT), lds, s, r);
    return hipGetLastError();
}

}  // namespace st
