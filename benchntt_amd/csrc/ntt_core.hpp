// ntt_core.hpp -- device building blocks of the forward constant-geometry NTT,
// shared by the launch-per-call kernel (ntt.hip) and the persistent
// work-queue kernel (queue.hpp: uneven batches in batch.hip, polynomial
// products in serve.hip).
//
// Reference network (ntt_scalar.c:582-627): log2(N) stages; at stage s
// (t = 2^s), for j in [0, N/2):
//     p = mulmod(tw[t + (j mod t)], in[j + N/2])
//     out[2j] = addmod(in[j], p),  out[2j+1] = submod(in[j], p).
// Each stage pairs the elements whose indices differ in the top bit and writes
// them to positions that differ in the bottom bit: the index rotates left by
// one.  Hence any k consecutive stages starting at stage S act independently
// on the 2^(L-k) groups of 2^k elements that share their low L-k index bits M
// (inputs at stride 2^(L-k)), and each group's outputs land contiguously at
// M*2^k.  Inside a group, local stage r is again a constant-geometry stage of
// size 2^k whose butterfly i uses
//     tw[2^(S+r) + ((M mod 2^S) << r) + (i mod 2^r)].
// The same split applies recursively inside a group: a workgroup holds G
// groups in LDS and runs register sub-passes of up to 2^ELOG points per
// thread (LDS -> VGPR -> LDS) until all k stages are done.
#pragma once

#include "kernels.hpp"

namespace bntt {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Device-memory accesses go through the global address space.  Pointers read
// out of descriptors are generic to the compiler, and its flat loads count on
// the LDS counter as well, so every wait for an LDS read would also wait for
// the outstanding device-memory loads.  (Every such pointer here is device or
// pinned host memory, never LDS or scratch.)
typedef __attribute__((address_space(1))) const v4u *gv4u_cptr;
typedef __attribute__((address_space(1))) v4u *gv4u_ptr;

__device__ __forceinline__ U128 ld16(const U128 *p)
{
    const v4u t = *(gv4u_cptr)(reinterpret_cast<const v4u *>(p));
    U128 r;
    r.v[0] = t.x;
    r.v[1] = t.y;
    r.v[2] = t.z;
    r.v[3] = t.w;
    return r;
}

__device__ __forceinline__ void st16(U128 *p, const U128 &a)
{
    v4u t;
    t.x = a.v[0];
    t.y = a.v[1];
    t.z = a.v[2];
    t.w = a.v[3];
    *(gv4u_ptr)(reinterpret_cast<v4u *>(p)) = t;
}

// A digit twiddle from a table in device memory (16-byte entries).
__device__ __forceinline__ R31 ldr31(const R31 *p)
{
    const v4u t = *(gv4u_cptr)(reinterpret_cast<const v4u *>(p));
    R31 r;
    r.d[0] = t.x;
    r.d[1] = t.y;
    r.d[2] = t.z;
    r.d[3] = t.w;
    return r;
}

constexpr int kMaxThreads = 512;
// Dynamic LDS of the largest launch: a 4096-point transform and its 4095
// staged twiddles, padded (the CU has 160 KiB).
constexpr size_t kMaxLdsBytes = 148 * 1024;
// dynamic LDS per workgroup that leaves room for three on a CU (160 KiB / 3,
// less the kernel's static LDS)
constexpr size_t kLds3Bytes = 53 * 1024;

// LDS placement.  Data element i of groups of 2^k elements sits at
//   lidx(i, k) = i + (i >> 3) + (i >> k):
// one pad slot after every 8 elements spreads the stride-2, -4 and -8
// write-backs of the register sub-passes over all banks, and one extra slot per
// group spreads the stride-2^k writes of the multi-pass loads (consecutive
// lanes take consecutive groups).  Staged twiddle entry e sits at tidx(e).
__device__ __forceinline__ int lidx(int i, int k) { return i + (i >> 3) + (i >> k); }
__device__ __forceinline__ int tidx(int e) { return e + (e >> 3); }
// Stages per pass of an np-pass (1..3) transform of 2^L points whose passes
// hold 2^k-point groups (elog <= k <= tile_log) and run register sub-passes of
// up to elog stages: the split with the fewest sub-passes (sum of
// ceil(k / elog)), then the smallest largest pass, then the most even, then
// the smaller passes first (pass 0 gathers the longest contiguous runs).
// False if none exists.
__host__ __device__ inline bool pass_split(int L, int np, int elog, int tile_log, int *ks)
{
    if (np == 1) {
        ks[0] = L;
        return L >= (elog < L ? elog : L) && L <= tile_log;
    }
    int best = 1 << 30;
    bool found = false;
    for (int k0 = elog; k0 <= tile_log; k0++) {
        for (int k1 = elog; k1 <= tile_log; k1++) {
            const int k2 = L - k0 - k1;
            if (np == 2 && k2 != 0) continue;
            if (np == 3 && (k2 < elog || k2 > tile_log)) continue;
            const int k2e = np == 3 ? k2 : k1;  // np = 2: two passes
            const int kmax = k0 > k1 ? (k0 > k2e ? k0 : k2e) : (k1 > k2e ? k1 : k2e);
            const int kmin = k0 < k1 ? (k0 < k2e ? k0 : k2e) : (k1 < k2e ? k1 : k2e);
            const int subs = (k0 + elog - 1) / elog + (k1 + elog - 1) / elog + (np == 3 ? (k2 + elog - 1) / elog : 0);
            const bool ascending = k0 <= k1 && (np == 2 || k1 <= k2);
            const int key = subs * 65536 + kmax * 256 + (kmax - kmin) * 2 + (ascending ? 0 : 1);
            if (key < best) {
                best = key;
                ks[0] = k0;
                ks[1] = k1;
                if (np == 3) ks[2] = k2;
                found = true;
            }
        }
    }
    return found;
}

__host__ __device__ inline size_t lds_data_slots(int k, int Glog)
{
    const size_t n = (size_t)1 << (Glog + k);
    return n + n / 8 + ((size_t)1 << Glog);
}

// The fast path's conditions.  Its product equals the reference's Barrett
// result for operands below 2^124.  When every input of a workgroup is at most
// q and every twiddle is below 2^124, every value the network produces is at
// most q (addmod and submod of values <= q, and products below q, stay <= q),
// so all its products qualify; otherwise the workgroup runs the literal
// reference formulas.
//   near_q:  a's top limb reaches q's (necessary for a > q; one compare, made
//            as values are loaded)
//   above_q: a > q exactly (run over a workgroup's tile only when some value
//            was near q)
//   tw_big:  a >= 2^124.
__device__ __forceinline__ bool near_q(const U128 &a, const U128 &q) { return a.v[3] >= q.v[3]; }
__device__ __forceinline__ bool above_q(const U128 &a, const U128 &q)
{
    uint32_t b;
    (void)subb(q.v[0], a.v[0], 0, b);
    (void)subb(q.v[1], a.v[1], b, b);
    (void)subb(q.v[2], a.v[2], b, b);
    (void)subb(q.v[3], a.v[3], b, b);
    return b != 0;
}
__device__ __forceinline__ bool tw_big(const U128 &a) { return (a.v[3] >> 28) != 0; }

// Whether a workgroup whose loads flagged near (any lane, near_q) holds a
// value above q: the exact test over the n values of its tile in LDS.
__device__ __forceinline__ bool tile_above_q(const U128 *lds, int n, int k, const U128 &q)
{
    bool big = false;
    for (int p = threadIdx.x; p < n; p += blockDim.x) big |= above_q(lds[lidx(p, k)], q);
    return __syncthreads_or(big);
}

// Where a register sub-pass finds its twiddles.
//   TW_LDS:    staged digit tables in LDS.  Stage table r holds, for the W
//              distinct group prefixes of the workgroup, the entries
//              tw[2^(S+r) + ((P0_start + rel) << r) + j], rel < W, j < 2^r, at
//              offset W*(2^r - 1) + (rel << r) + j (placed at tidx()).
//   TW_GLOBAL: a digit copy of the whole table in device memory, indexed like
//              the table itself (cached in L1/L2).
//   TW_SLOW:   the caller's table, literal reference arithmetic.
//   TW_MONT:   like TW_GLOBAL, but the digit table holds tw * 2^124 mod q and
//              the arithmetic is the Montgomery product with lazy butterflies
//              (values below 4q; polynomial-multiply path, any odd q < 2^122).
//   TW_MONT_LDS: TW_MONT arithmetic with TW_LDS staged tables.
//   TW_G128:   the caller's table, entries converted to digits in registers,
//              fast arithmetic (stages left unstaged by a TW_LDS workgroup).
enum TwMode { TW_SLOW = 0, TW_LDS = 1, TW_GLOBAL = 2, TW_MONT = 3, TW_MONT_LDS = 4, TW_G128 = 5, TW_MONT_S = 6, TW_MONT_S_LDS = 7 };

// Montgomery modes for a modulus whose 31-bit digits 1 and 2 are zero
__host__ __device__ constexpr bool tw_sparse(int mode) { return mode == TW_MONT_S || mode == TW_MONT_S_LDS; }
__host__ __device__ constexpr bool tw_in_lds(int mode) { return mode == TW_LDS || mode == TW_MONT_LDS || mode == TW_MONT_S_LDS; }
__host__ __device__ constexpr bool tw_mont(int mode) { return mode == TW_MONT || mode == TW_MONT_LDS || tw_sparse(mode); }

struct TwView {
    const R31 *lds;     // TW_LDS tables
    const R31 *gr31;    // TW_GLOBAL table
    const U128 *glob;   // the caller's table (TW_SLOW)
    uint32_t W;         // distinct prefixes staged (TW_LDS)
};

// Digit twiddles of one register sub-pass, fetched before its LDS reads so that
// their latency overlaps them: stage rp uses t[2^rp - 1 + j], j < 2^rp.
template <int KP, int MODE>
__device__ __forceinline__ void fetch_twiddles(R31 (&t)[(1 << KP) - 1], const TwView &tv, int S, int r0,
                                               uint32_t rel, uint32_t P0, uint32_t Mlow)
{
    const uint32_t P = (P0 << r0) | Mlow;
#pragma unroll
    for (int rp = 0; rp < KP; rp++) {
        const int r = r0 + rp;
        if constexpr (tw_in_lds(MODE)) {
            const int tb = (int)(tv.W * ((1u << r) - 1u) + (rel << r) + (Mlow << rp));
#pragma unroll
            for (int j = 0; j < (1 << rp); j++) t[(1 << rp) - 1 + j] = tv.lds[tidx(tb + j)];
        } else if constexpr (MODE == TW_G128) {
            const U128 *tb = tv.glob + (1u << (S + r)) + (P << rp);
#pragma unroll
            for (int j = 0; j < (1 << rp); j++) t[(1 << rp) - 1 + j] = to_r31(ld16(tb + j));
        } else {
            const R31 *tb = tv.gr31 + (1u << (S + r)) + (P << rp);
#pragma unroll
            for (int j = 0; j < (1 << rp); j++) t[(1 << rp) - 1 + j] = ldr31(tb + j);
        }
    }
}

// KP stages of the constant-geometry network on 2^KP values in registers with
// fetched digit twiddles (fast exact reduction; the modulus is of the fast
// form, so the add and subtract take its upper limbs as constants).
template <int KP>
__device__ __forceinline__ void reg_network_fast(U128 (&v)[1 << KP], const R31 (&t)[(1 << KP) - 1], const Mod &md)
{
    constexpr int H = 1 << (KP - 1);
#pragma unroll
    for (int rp = 0; rp < KP; rp++) {
        U128 w[1 << KP];
#pragma unroll
        for (int i = 0; i < H; i++) {
            const U128 p = mulmod_r31(v[i + H], t[(1 << rp) - 1 + (i & ((1 << rp) - 1))], md.c);
            w[2 * i] = addmod_pm(v[i], p, md.m.v[0]);
            w[2 * i + 1] = submod_pm(v[i], p, md.m.v[0]);
        }
#pragma unroll
        for (int i = 0; i < 2 * H; i++) v[i] = w[i];
    }
}

// The same stages with Montgomery-form digit twiddles, montmul(v, tw*R) =
// v*tw, and lazy reduction (Harvey's butterfly): values stay below 4q; the
// even input is brought below 2q, the product is below 2q, and the outputs
// x + p and x + 2q - p are below 4q.  The caller makes the results canonical.
template <int KP, bool SPARSE>
__device__ __forceinline__ void reg_network_mont(U128 (&v)[1 << KP], const R31 (&t)[(1 << KP) - 1], const Mod &md)
{
    constexpr int H = 1 << (KP - 1);
#pragma unroll
    for (int rp = 0; rp < KP; rp++) {
        U128 w[1 << KP];
#pragma unroll
        for (int i = 0; i < H; i++) {
            const U128 x = sub_if_ge(v[i], md.mq.ql2);
            const U128 p = montmul_r31_lazy<SPARSE>(v[i + H], t[(1 << rp) - 1 + (i & ((1 << rp) - 1))], md.mq);
            w[2 * i] = add128l(x, p);
            w[2 * i + 1] = add128l(x, sub128l(md.mq.ql2, p));
        }
#pragma unroll
        for (int i = 0; i < 2 * H; i++) v[i] = w[i];
    }
}

// The same stages with the caller's table and the literal reference formulas
// (used when some input of the workgroup is above q or a twiddle >= 2^124, or the modulus is not
// of the fast form).  Local stages r0 .. r0+KP-1 of a group network that starts
// at global stage S; P0: the group's prefix (low S bits of its M); Mlow: low r0
// bits of the sub-group index.
template <int KP, bool PM>
__device__ __forceinline__ void reg_network_slow(U128 (&v)[1 << KP], const TwView &tv, int S, int r0, uint32_t P0,
                                                 uint32_t Mlow, const Mod &md)
{
    constexpr int H = 1 << (KP - 1);
    const uint32_t P = (P0 << r0) | Mlow;
#pragma unroll
    for (int rp = 0; rp < KP; rp++) {
        const int r = r0 + rp;
        const U128 *tb = tv.glob + (1u << (S + r)) + (P << rp);
        U128 t[H];
#pragma unroll
        for (int j = 0; j < (1 << rp) && j < H; j++) t[j] = ld16(tb + j);
        U128 w[1 << KP];
#pragma unroll
        for (int i = 0; i < H; i++) {
            const U128 p = mulmod<PM>(t[i & ((1 << rp) - 1)], v[i + H], md);
            w[2 * i] = addmod(v[i], p, md);
            w[2 * i + 1] = submod(v[i], p, md);
        }
#pragma unroll
        for (int i = 0; i < 2 * H; i++) v[i] = w[i];
    }
}

// One register sub-pass of KP stages over the G groups of 2^k values in LDS.
// Each thread owns 2^(ELOG-KP) sub-groups.  Group g of the tile is group
// mbase + (g & pmask) of its transform (pmask < G - 1 when the tile holds the
// same groups of two transforms, one after the other).
template <int ELOG, int KP, int MODE, bool PM>
__device__ __forceinline__ void lds_subpass(U128 *lds, const TwView &tv, int k, int S, int r0, uint32_t mbase,
                                            uint32_t pmask, const Mod &md)
{
    constexpr int Q = 1 << (ELOG - KP);
    constexpr int R = 1 << KP;
    constexpr int NT = MODE == TW_SLOW ? 1 : (1 << KP) - 1;
    const int T = blockDim.x;
    const int sub_log = k - KP;  // sub-groups per group = 2^sub_log
    U128 v[Q][R];
    R31 t[Q][NT];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = threadIdx.x + q * T;
        const int g = sg >> sub_log;
        const int Mp = sg & ((1 << sub_log) - 1);
        if constexpr (MODE != TW_SLOW) {
            const uint32_t gp = (uint32_t)g & pmask;
            const uint32_t P0 = (mbase + gp) & ((1u << S) - 1u);
            const uint32_t Mlow = (uint32_t)Mp & ((1u << r0) - 1u);
            fetch_twiddles<KP, MODE>(t[q], tv, S, r0, gp & (tv.W - 1u), P0, Mlow);
        }
        const int src = (g << k) + Mp;
        if (sub_log >= 3) {
            // the sub-group's values are 2^sub_log apart inside one group:
            // their LDS slots are a fixed stride apart
            const int base = lidx(src, k), stride = (1 << sub_log) + (1 << (sub_log - 3));
#pragma unroll
            for (int j = 0; j < R; j++) v[q][j] = lds[base + j * stride];
        } else {
#pragma unroll
            for (int j = 0; j < R; j++) v[q][j] = lds[lidx(src + (j << sub_log), k)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = threadIdx.x + q * T;
        const int g = sg >> sub_log;
        const int Mp = sg & ((1 << sub_log) - 1);
        if constexpr (tw_mont(MODE)) {
            reg_network_mont<KP, tw_sparse(MODE)>(v[q], t[q], md);
        } else if constexpr (MODE != TW_SLOW) {
            reg_network_fast<KP>(v[q], t[q], md);
        } else {
            const uint32_t P0 = (mbase + ((uint32_t)g & pmask)) & ((1u << S) - 1u);
            const uint32_t Mlow = (uint32_t)Mp & ((1u << r0) - 1u);
            reg_network_slow<KP, PM>(v[q], tv, S, r0, P0, Mlow, md);
        }
        const int dst = (g << k) + (Mp << KP);
        if (KP == 3 && k >= 3) {
            // 8 consecutive values of one aligned block of 8: consecutive slots
            const int base = lidx(dst, k);
#pragma unroll
            for (int j = 0; j < R; j++) lds[base + j] = v[q][j];
        } else {
#pragma unroll
            for (int j = 0; j < R; j++) lds[lidx(dst + j, k)] = v[q][j];
        }
    }
    __syncthreads();
}

// A network whose tables are staged in LDS (MODE = TW_LDS or TW_MONT_LDS)
// only for its first klds stages runs the sub-passes that reach beyond them
// with FALLBACK (TW_G128 or TW_GLOBAL, or TW_MONT for TW_MONT_LDS); FALLBACK
// = -1: every sub-pass uses MODE.
template <int ELOG, int MODE, bool PM, int FALLBACK>
__device__ __forceinline__ void run_subpass(U128 *lds, const TwView &tv, int k, int S, int r0, int kp, bool lds_tw,
                                            uint32_t mbase, uint32_t pmask, const Mod &md)
{
    if constexpr (FALLBACK >= 0) {
        if (!lds_tw) {
            if (kp == ELOG) {
                lds_subpass<ELOG, ELOG, FALLBACK, PM>(lds, tv, k, S, r0, mbase, pmask, md);
            } else if constexpr (ELOG >= 2) {
                if (kp == 1) {
                    lds_subpass<ELOG, 1, FALLBACK, PM>(lds, tv, k, S, r0, mbase, pmask, md);
                } else if constexpr (ELOG >= 3) {
                    if (kp == 2) lds_subpass<ELOG, 2, FALLBACK, PM>(lds, tv, k, S, r0, mbase, pmask, md);
                }
            }
            return;
        }
    }
    if (kp == ELOG) {
        lds_subpass<ELOG, ELOG, MODE, PM>(lds, tv, k, S, r0, mbase, pmask, md);
    } else if constexpr (ELOG >= 2) {
        if (kp == 1) {
            lds_subpass<ELOG, 1, MODE, PM>(lds, tv, k, S, r0, mbase, pmask, md);
        } else if constexpr (ELOG >= 3) {
            if (kp == 2) lds_subpass<ELOG, 2, MODE, PM>(lds, tv, k, S, r0, mbase, pmask, md);
        }
    }
}

template <int ELOG, int MODE, bool PM, int FALLBACK = -1>
__device__ __forceinline__ void run_group_network(U128 *lds, const TwView &tv, int k, int S, uint32_t mbase,
                                                  const Mod &md, int klds = 32, uint32_t pmask = ~0u)
{
    int r0 = 0;
    while (r0 < k) {
        const int kp = (k - r0) < ELOG ? (k - r0) : ELOG;
        run_subpass<ELOG, MODE, PM, FALLBACK>(lds, tv, k, S, r0, kp, r0 + kp <= klds, mbase, pmask, md);
        r0 += kp;
    }
}

// Stage, as digits, the twiddles that stages [S, S+klds) use for W = 2^wlog
// consecutive group prefixes starting at P0start: stage table r (2^r entries
// per prefix) sits at LDS offset W*(2^r - 1).  Returns whether any entry of
// stages [S, S+k) is >= 2^124.
__device__ __forceinline__ bool stage_twiddles(R31 *tws, const U128 *__restrict__ tw, int S, int k, int wlog,
                                               uint32_t P0start, int klds)
{
    bool big = false;
    const int ntw = (int)(((1u << k) - 1u) << wlog);
    const int nstaged = (int)(((1u << (klds < k ? klds : k)) - 1u) << wlog);  // stages r < klds come first
    // four loads of a thread in flight at once (an entry past the tables reads
    // entry 1 instead and is not kept)
    constexpr int U = 4;
    const int T = blockDim.x;
    if (S == 0) {
        // the first stages of a transform (one prefix): the staged tables are
        // the table itself from entry 1 on
        for (int e0 = threadIdx.x; e0 < ntw; e0 += U * T) {
            U128 a[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * T;
                a[u] = ld16(tw + 1 + (e < ntw ? e : 0));
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * T;
                if (e < ntw) {
                    big |= tw_big(a[u]);
                    if (e < nstaged) tws[tidx(e)] = to_r31(a[u]);
                }
            }
        }
        return big;
    }
    for (int e0 = threadIdx.x; e0 < ntw; e0 += U * T) {
        U128 a[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // e = W*(2^r - 1) + rel*2^r + j  ->  stage r, offset within the stage table
            const int e = e0 + u * T < ntw ? e0 + u * T : 0;
            const uint32_t v = ((uint32_t)e >> wlog) + 1u;  // in [2^r, 2^(r+1))
            const int r = 31 - __builtin_clz(v);
            const uint32_t off = (uint32_t)e - (((1u << r) - 1u) << wlog);
            a[u] = ld16(tw + (1u << (S + r)) + (P0start << r) + off);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + u * T;
            if (e < ntw) {
                big |= tw_big(a[u]);
                if (e < nstaged) tws[tidx(e)] = to_r31(a[u]);
            }
        }
    }
    return big;
}

// The same staging from a table already in digits (no check: the caller
// knows whether the table's entries are below 2^124).
__device__ __forceinline__ void stage_twiddles_r31(R31 *tws, const R31 *__restrict__ twr, int S, int klds, int wlog,
                                                   uint32_t P0start)
{
    const int ntw = (int)(((1u << klds) - 1u) << wlog);  // stages [S, S + klds)
    if (S == 0) {
        for (int e = threadIdx.x; e < ntw; e += blockDim.x) tws[tidx(e)] = ldr31(twr + 1 + e);
        return;
    }
    for (int e = threadIdx.x; e < ntw; e += blockDim.x) {
        const uint32_t u = ((uint32_t)e >> wlog) + 1u;  // in [2^r, 2^(r+1))
        const int r = 31 - __builtin_clz(u);
        const uint32_t off = (uint32_t)e - (((1u << r) - 1u) << wlog);
        tws[tidx(e)] = ldr31(twr + (1u << (S + r)) + (P0start << r) + off);
    }
}

// Digit slots of the staged tables of k stages for 2^wlog prefixes.
__host__ __device__ inline size_t tw_slots(int k, int wlog)
{
    const size_t ntw = (((size_t)1 << k) - 1) << wlog;
    return ntw + ntw / 8 + 1;
}

// LDS bytes of one workgroup: the data, then the digit tables of the first
// klds stages.
__host__ __device__ inline size_t pass_lds_bytes(int k, int S, int Glog, int klds)
{
    const int wlog = Glog < S ? Glog : S;
    const size_t ntw = ((size_t)1 << wlog) * (((size_t)1 << (klds < k ? klds : k)) - 1);
    return sizeof(U128) * lds_data_slots(k, Glog) + sizeof(R31) * (ntw + ntw / 8 + 1);
}

}  // namespace bntt
