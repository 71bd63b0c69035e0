// queue.hpp -- the persistent work-queue kernel and its work items, shared by
// the uneven-batch planner (batch.hip, reference arithmetic) and the
// polynomial-multiply server (serve.hip, exact Montgomery arithmetic).
//
// A launch processes one list of work items: a grid sized to the resident
// capacity of the chip (at most one workgroup per item) whose workgroups take
// items from an atomic counter until the list is exhausted; the last draw
// rearms the counter for the next launch.  Kinds of launch:
//   QK_FAST  reference arithmetic on its fast path only (modulus 2^124 - c):
//            an item that needs the literal formulas (a value or twiddle
//            >= 2^124, or a product operand >= 2^124) is left untouched and
//            its index appended to the launch's punt list
//   QK_SLOW  the literal reference formulas for every item it takes: the
//            punt list of the fast launch before it, or a whole list when the
//            modulus is not of the fast form
//   QK_FAST8 QK_FAST for launches whose transform items all have 2048-value
//            tiles (256 threads x radix-8 register sub-passes) and read their
//            twiddle digits from device memory: each item runs code
//            specialised for its stage count K (do_ntt_item8<K>), whose LDS
//            and twiddle addresses are fixed per item
//   QK_MONT  exact Montgomery arithmetic (polynomial-multiply server);
//            QK_MONT_S the same for a modulus whose 31-bit digits 1 and 2
//            are zero (a shorter reduction); QK_MONT8 / QK_MONT_S8 the same
//            with the per-stage-count code of QK_FAST8
// Keeping the literal formulas out of the fast kernel keeps its registers
// free of spills.  Kinds of item:
//   ITEM_NTT   stages [S, S+k) of the forward network for G groups (several
//              whole transforms when k = log2 N, else blocks of one transform)
//   ITEM_MUL   2048 elements of c = a * b mod q
// With reference arithmetic a transform item may hold two transforms A and B
// of one size and one table (IT_PAIR): groups [0, H) of the tile are A's and
// [H, 2H) are B's (H = G/2), the same group indices of each (blocks of the
// last pass of multi-pass transforms, whose earlier passes run unpaired) or
// the A and B of up to H pairs (whole transforms; their descriptors are the H
// A's, then the H B's).  The two share their twiddles.
// IT_PROD on such an item's last pass also writes the pointwise product
// c = y_A * y_B mod q (c: the A descriptor's prod) from the tile, so the
// product reads nothing back from device memory.
// With Montgomery arithmetic a transform item may also fuse the pointwise
// steps of a negacyclic product into its first load and its last store:
//   IT_GATHER_MUL     the first pass reads x[i] = A[brv(i)] * B[brv(i)] * R^-1
//   IT_GATHER         the first pass reads x[i] = A[brv(i)] * R^-1
//   IT_SCATTER_SCALE  the last pass writes C[brv(i)] = y[i] * S[i] * R^-1
//   IT_CANON          the last pass writes y canonical (values in [0, q))
// (values between the Montgomery steps are kept below 4q; C is canonical)
// where brv reverses the log2 N index bits of a transform.  IT_CONTIG marks a
// packed item whose transforms lie back to back in x and in y: it is read and
// written through the first descriptor's pointers.
#pragma once

#include "ntt_core.hpp"

namespace bntt {

constexpr int kQueueThreads = 256;
constexpr int kQueueMaxElog = 3;
constexpr int kQueueTile = kQueueThreads << kQueueMaxElog;  // 2048 values
// padded data tile, then (when a launch reserves room) staged digit twiddles
constexpr size_t kQueueMaxLds = 80 * 1024;
// dynamic LDS that keeps three workgroups on a CU (160 KiB / 3, less the
// static arrays of the kernel, up to 6.2 KiB)
constexpr size_t kQueueLds3 = 46 * 1024;

enum ItemKind : uint8_t { ITEM_NTT = 0, ITEM_MUL = 1 };
enum ItemFlags : uint8_t {
    IT_GATHER_MUL = 1,
    IT_SCATTER_SCALE = 2,
    IT_CONTIG = 4,
    IT_GATHER = 8,
    IT_CANON = 16,
    IT_PAIR = 32,
    IT_PROD = 64
};

struct DevDesc {
    const U128 *x;
    U128 *y;
    U128 *tmp;
    const U128 *tw;
    const R31 *twr;          // digit copy of tw (reference arithmetic) or tw*R digits (Montgomery)
    const uint32_t *twflag;  // nonzero when some entry of tw (past entry 0) is >= 2^124
    const U128 *x2;          // IT_GATHER_MUL: second factor B (x holds A)
    const R31 *scale;        // IT_SCATTER_SCALE: S in digits
    U128 *prod;              // IT_PROD (A's descriptor): c = y_A * y_B
};

struct DevMul {
    const U128 *a;
    const U128 *b;
    U128 *c;
    uint64_t n;
};

struct WorkItem {
    uint8_t kind, L, S, k;
    uint8_t Glog, elog, pass, np;
    uint32_t desc;    // descriptor / product / point index
    uint32_t nvalid;  // groups (NTT) or elements (products)
    uint64_t first;   // first group of the pass (NTT multi-pass) or element offset (products)
    uint8_t flags;    // ItemFlags (Montgomery transform items)
    uint8_t pad[7];
};
static_assert(sizeof(WorkItem) == 32, "work item layout");

struct QueueCtl {
    uint32_t ticket;   // dynamic draws of the running launch; 0 between launches
    uint32_t readers;  // slow launches over a punt list: workgroups that have read its count
    uint32_t pad[2];
};

enum QueueKind : int { QK_FAST = 0, QK_SLOW = 1, QK_MONT = 2, QK_MONT_S = 3, QK_FAST8 = 4, QK_MONT8 = 5, QK_MONT_S8 = 6 };
__host__ __device__ constexpr bool qk_mont(int kind)
{
    return kind == QK_MONT || kind == QK_MONT_S || kind == QK_MONT8 || kind == QK_MONT_S8;
}
// the shorter Montgomery reduction (q's 31-bit digits 1 and 2 are zero)
__host__ __device__ constexpr bool qk_sparse(int kind) { return kind == QK_MONT_S || kind == QK_MONT_S8; }
// fast-path reference arithmetic: items that need the literal formulas are punted
__host__ __device__ constexpr bool qk_fast(int kind) { return kind == QK_FAST || kind == QK_FAST8; }
// launches of 2048-value radix-8 items run code specialised per stage count
__host__ __device__ constexpr bool qk_item8(int kind)
{
    return kind == QK_FAST8 || kind == QK_MONT8 || kind == QK_MONT_S8;
}

// Items a fast launch left to the literal formulas: count is cleared before
// each execution, idx holds up to one entry per item of the launch's list.
struct PuntList {
    uint32_t *count;
    uint32_t *idx;
};

// Conversion of twiddle tables to digits.
struct TwJob {
    const U128 *src;
    R31 *dst;
    uint32_t *flag;
    uint64_t n;
};
struct TwChunk {
    uint32_t job;
    uint32_t count;
    uint64_t off;
};
constexpr int kTwChunk = 4096;

// klds > 0: the tables of the first klds stages are staged in LDS, the later
// stages read the digit table in device memory.
template <int ELOG, int KIND>
__device__ __forceinline__ void run_network_elog(U128 *lds, const TwView &tv, int k, int S, uint32_t mbase, int klds,
                                                 uint32_t pmask, const Mod &md)
{
    if constexpr (qk_mont(KIND)) {
        constexpr int M = qk_sparse(KIND) ? TW_MONT_S : TW_MONT;
        constexpr int ML = qk_sparse(KIND) ? TW_MONT_S_LDS : TW_MONT_LDS;
        if (klds > 0)
            run_group_network<ELOG, ML, true, M>(lds, tv, k, S, mbase, md, klds, pmask);
        else
            run_group_network<ELOG, M, true>(lds, tv, k, S, mbase, md, 32, pmask);
    } else if constexpr (KIND == QK_SLOW) {
        // the literal Barrett product whatever the modulus
        run_group_network<ELOG, TW_SLOW, false>(lds, tv, k, S, mbase, md, 32, pmask);
    } else {
        if (klds > 0)
            run_group_network<ELOG, TW_LDS, true, TW_GLOBAL>(lds, tv, k, S, mbase, md, klds, pmask);
        else
            run_group_network<ELOG, TW_GLOBAL, true>(lds, tv, k, S, mbase, md, 32, pmask);
    }
}

// LDS layout of a launch: the data tile in U128 slots [0, tw_off), then room
// for tw_cap staged digit twiddles.  The staged tables stay valid across the
// items of a workgroup until an item needs different ones (key).
struct TwArea {
    uint32_t tw_off;
    uint32_t tw_cap;
};

// log2 of the groups of one transform in an item's tile (the tile's Glog,
// less one for an item holding two transforms).
__host__ __device__ inline int item_plog(const WorkItem &it) { return it.Glog - ((it.flags & IT_PAIR) ? 1 : 0); }

// Twiddle-table prefixes staged for a transform item (see do_ntt_item); plog
// from item_plog.
__host__ __device__ inline int item_tw_wlog(int L, int k, int S, int plog)
{
    return L == k ? 0 : (plog < S ? plog : S);
}

// Whether some kept value of a tile is above q (the exact test, run only in a
// workgroup whose loads flagged a value near q): groups g with
// (g & pmask) < nvalid are kept.
__device__ __forceinline__ bool tile_above_q_kept(const U128 *lds, int k, int Glog, uint32_t pmask, int nvalid,
                                                  const U128 &q)
{
    bool big = false;
    for (int p = threadIdx.x; p < (1 << (Glog + k)); p += blockDim.x)
        if ((int)((uint32_t)(p >> k) & pmask) < nvalid) big |= above_q(lds[lidx(p, k)], q);
    return __syncthreads_or(big);
}

// The source and destination of an item's pass for one descriptor.
__device__ __forceinline__ void pass_buffers(const WorkItem &it, const DevDesc &d, const U128 *&src, U128 *&dst)
{
    if (it.np == 1) {
        src = d.x;
        dst = d.y;
    } else {
        const int p = it.pass, np = it.np;
        dst = ((np - 1 - p) % 2 == 0) ? d.y : d.tmp;
        src = (p == 0) ? d.x : (((np - p) % 2 == 0) ? d.y : d.tmp);
    }
}

// Layout of the intermediate of a two-pass transform: the first pass stores
// it in the second pass's group order (position p of its group M at
// (p << (L - k)) + M, consecutive groups side by side), so that each group of
// the second pass is one contiguous block (its group M' at M' << k').  Reads
// and writes then move runs of at least a tile's group count.  (Three-pass
// transforms keep every pass's outputs in network order: group M at M << k.)
__host__ __device__ inline bool first_of_two(const WorkItem &it) { return it.np == 2 && it.pass == 0; }
__host__ __device__ inline bool second_of_two(const WorkItem &it) { return it.np == 2 && it.pass == 1; }

// Set up the per-transform buffers of an item of whole transforms (one per
// group): s_x[g], s_y[g] for the tile's groups, s_c[p] for the products of a
// paired item, s_x2 for Montgomery gathers.  Ends with a barrier.
template <bool MONT>
__device__ __forceinline__ void item_buffers(const U128 **s_x, U128 **s_y, const U128 **s_x2, U128 **s_c,
                                             const WorkItem &it, const DevDesc *__restrict__ descs)
{
    const int nvalid = (int)it.nvalid;
    const int t = threadIdx.x;
    if (!MONT && (it.flags & IT_PAIR)) {
        // descriptors: the nvalid A's, then the nvalid B's; tile groups [0, H)
        // are the A's and [H, 2H) the B's
        const int H = 1 << (it.Glog - 1);
        if (t < nvalid) {
            const DevDesc &da = descs[it.desc + t];
            const DevDesc &db = descs[it.desc + nvalid + t];
            s_x[t] = da.x;
            s_y[t] = da.y;
            s_c[t] = da.prod;
            s_x[H + t] = db.x;
            s_y[H + t] = db.y;
        }
    } else if (t < nvalid) {
        const DevDesc &d = descs[it.desc + t];
        s_x[t] = d.x;
        s_y[t] = d.y;
        if constexpr (MONT) s_x2[t] = d.x2;
    }
    __syncthreads();
}

// Layout of a launch over items whose data tiles need at most data_slots and
// whose full staged tables need tw_needs[i] digit slots: the twiddle area gets
// what `budget` bytes of dynamic LDS leave beside the data, at most the
// largest need (an item stages the tables of as many of its first stages as
// fit).  Returns the dynamic LDS bytes.
inline size_t plan_tw_area(size_t data_slots, const size_t *tw_needs, size_t n, size_t budget, TwArea *area)
{
    size_t need = 0;
    for (size_t i = 0; i < n; i++) need = tw_needs[i] > need ? tw_needs[i] : need;
    const size_t data = sizeof(U128) * data_slots;
    size_t cap = budget > data ? (budget - data) / sizeof(R31) : 0;
    if (cap > need) cap = need;
    area->tw_off = (uint32_t)data_slots;
    area->tw_cap = (uint32_t)cap;
    return data + sizeof(R31) * cap;
}

// a value below 4q brought into [0, q)
__device__ __forceinline__ U128 canonical(const U128 &a, const MontQ &mq)
{
    return sub_if_ge(sub_if_ge(a, mq.ql2), mq.ql);
}

// A * B * R^-1 for forward outputs A, B below 4q (lazy butterflies): B is
// brought below q, the product is below 2q (a valid network input).
template <bool SPARSE>
__device__ __forceinline__ U128 gather_product(const U128 &a, const U128 &b, const MontQ &mq)
{
    return montmul_r31_lazy<SPARSE>(a, to_r31(canonical(b, mq)), mq);
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, uint32_t log2n)
{
    return log2n ? (__brev(i) >> (32u - log2n)) : 0u;
}

// HBM -> LDS for one transform item, E values per thread.  All of a thread's
// loads are issued before the first is used (an element past the item's
// groups reads its first group instead and is not kept), so they are in
// flight together rather than one at a time.  Returns whether a kept value is
// near q (reference arithmetic).
template <int E, int KIND>
__device__ __forceinline__ bool load_tile(U128 *lds, const U128 *const *s_x, const U128 *const *s_x2,
                                          const WorkItem &it, const DevDesc &d0, const U128 *src, const U128 *srcB,
                                          const Mod &md)
{
    constexpr bool MONT = qk_mont(KIND), SPARSE = qk_sparse(KIND);
    const int L = it.L, k = it.k, Glog = it.Glog, mlog = L - k, nvalid = (int)it.nvalid;
    const int G = 1 << Glog;
    // groups of one transform in the tile (a paired item holds two)
    const int plog = MONT ? Glog : item_plog(it);
    const uint32_t pmask = (1u << plog) - 1u;
    const bool gather = MONT && (it.flags & (IT_GATHER_MUL | IT_GATHER)) && it.pass == 0;
    const bool gmul = (it.flags & IT_GATHER_MUL) != 0;
    const bool contig = !MONT && (it.flags & IT_CONTIG);
    const U128 one{{1u, 0u, 0u, 0u}};
    U128 a[E], b[E];
    bool big = false;
    if (contig) {
        // transforms back to back: one contiguous range
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            a[j] = ld16(d0.x + ((e >> k) < nvalid ? e : 0));
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            if ((e >> k) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx(e, k)] = a[j];
            }
        }
    } else if (mlog == 0) {
        // whole transforms, one per group, each with its own buffers
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k, ti = e & ((1 << k) - 1);
            const int gv = (int)((uint32_t)g & pmask) < nvalid ? g : 0;
            a[j] = ld16(s_x[gv] + ti);
            if (gather && gmul) b[j] = ld16(s_x2[gv] + ti);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k, ti = e & ((1 << k) - 1);
            if ((int)((uint32_t)g & pmask) < nvalid) {
                if (gather) {
                    // network input brv(j) is A[j] * B[j]: contiguous loads,
                    // bit-reversed placement in LDS
                    lds[lidx((g << k) + (int)bitrev((uint32_t)ti, (uint32_t)k), k)] =
                        gather_product<SPARSE>(a[j], gmul ? b[j] : one, md.mq);
                } else {
                    big |= near_q(a[j], md.m);
                    lds[lidx(e, k)] = a[j];
                }
            }
        }
    } else if (gather) {
        // group g's inputs (ti << mlog) + M, bit-reversed, are the contiguous
        // block brv(M) << k .. + 2^k of A and B (in the order brv(ti)): load
        // the block, place each value at its ti
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k, jj = e & ((1 << k) - 1);
            const uint32_t gv = (uint32_t)(g < nvalid ? g : 0);
            const uint32_t blk = bitrev((uint32_t)it.first + gv, (uint32_t)mlog) << k;
            a[j] = ld16(d0.x + blk + jj);
            if (gmul) b[j] = ld16(d0.x2 + blk + jj);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k, jj = e & ((1 << k) - 1);
            if (g < nvalid)
                lds[lidx((g << k) + (int)bitrev((uint32_t)jj, (uint32_t)k), k)] =
                    gather_product<SPARSE>(a[j], gmul ? b[j] : one, md.mq);
        }
    } else if (second_of_two(it)) {
        // the first pass stored each group of this pass contiguously (group M
        // at M << k): the tile's groups (of A, then of B in a paired item) are
        // contiguous blocks
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k;
            const uint32_t gp = (uint32_t)g & pmask;
            const uint32_t gv = (int)gp < nvalid ? gp : 0u;
            const U128 *sb = (g >> plog) ? srcB : src;
            a[j] = ld16(sb + (((uint32_t)it.first + gv) << k) + (uint32_t)(e & ((1 << k) - 1)));
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            if ((int)((uint32_t)(e >> k) & pmask) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx(e, k)] = a[j];
            }
        }
    } else {
        // blocks of one transform (of A or B in a paired item): element ti of
        // group g at (ti << mlog) + M (indices within one transform, at most
        // 2^24 points, fit 32 bits)
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e & (G - 1), ti = e >> Glog;
            const uint32_t gp = (uint32_t)g & pmask;
            const uint32_t gv = (int)gp < nvalid ? gp : 0u;
            const U128 *sb = (g >> plog) ? srcB : src;
            a[j] = ld16(sb + ((uint32_t)ti << mlog) + (uint32_t)it.first + gv);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e & (G - 1), ti = e >> Glog;
            if ((int)((uint32_t)g & pmask) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx((g << k) + ti, k)] = a[j];
            }
        }
    }
    return big;
}

// The pointwise product of two network outputs at one position: c = A * B mod
// q, the reference's vector_mul (blas_scalar.c:594-600) of the two transforms'
// outputs.  Fast launches take the exact radix-2^31 product (every value is
// at most q there); slow launches the literal formula.
template <int KIND>
__device__ __forceinline__ U128 pair_product(const U128 &a, const U128 &b, const Mod &md)
{
    if constexpr (qk_fast(KIND))
        return mulmod_r31(b, to_r31(a), md.c);
    else
        return mulmod_ref(a, b, md.m, md.mu);
}

// One transform item: load, network, store.  Returns false when a fast
// launch leaves the item to the literal formulas (nothing written).
template <int KIND>
__device__ bool do_ntt_item(U128 *lds, const U128 **s_x, U128 **s_y, const U128 **s_x2, U128 **s_c,
                            uint64_t *s_twkey, const WorkItem &it, const DevDesc *__restrict__ descs, TwArea area,
                            const Mod &md)
{
    constexpr bool MONT = qk_mont(KIND), SPARSE = qk_sparse(KIND);
    const int L = it.L, S = it.S, k = it.k, Glog = it.Glog, elog = it.elog;
    const int E = 1 << elog;
    const int mlog = L - k;
    const int nvalid = (int)it.nvalid;
    bool big = false;
    const DevDesc d0 = descs[it.desc];
    const bool pair = !MONT && (it.flags & IT_PAIR);
    const int plog = MONT ? Glog : item_plog(it);
    const uint32_t pmask = (1u << plog) - 1u;
    const int H = 1 << plog;
    const bool prod = pair && (it.flags & IT_PROD) && it.pass + 1 == it.np;
    const U128 *src, *srcB = nullptr;
    U128 *dst, *dstB = nullptr;
    pass_buffers(it, d0, src, dst);
    if (pair && mlog > 0) pass_buffers(it, descs[it.desc + 1], srcB, dstB);
    // fused pointwise steps (Montgomery products only); a plain gather is a
    // product with B = 1
    const bool gather = MONT && (it.flags & (IT_GATHER_MUL | IT_GATHER)) && it.pass == 0;
    const bool scatter = MONT && (it.flags & IT_SCATTER_SCALE) && it.pass + 1 == it.np;
    const bool canon = MONT && (it.flags & IT_CANON) && it.pass + 1 == it.np;
    const bool contig = !MONT && (it.flags & IT_CONTIG);
    (void)gather;
    if (mlog == 0 && !contig) item_buffers<MONT>(s_x, s_y, s_x2, s_c, it, descs);
    // HBM -> LDS
    switch (elog) {
    case 1:
        big = load_tile<2, KIND>(lds, s_x, s_x2, it, d0, src, srcB, md);
        break;
    case 2:
        big = load_tile<4, KIND>(lds, s_x, s_x2, it, d0, src, srcB, md);
        break;
    default:
        big = load_tile<8, KIND>(lds, s_x, s_x2, it, d0, src, srcB, md);
        break;
    }
    const uint32_t mbase = (uint32_t)it.first;
    if constexpr (qk_fast(KIND)) {
        // a large twiddle, or a value above q (decided exactly in the rare
        // workgroups where some value is near q): the literal formulas.  The
        // vote is also the barrier after the tile writes.
        const bool twb = *(const __attribute__((address_space(1))) uint32_t *)d0.twflag != 0;  // one flag per item
        if (__syncthreads_or(big || twb) && (twb || tile_above_q_kept(lds, k, Glog, pmask, nvalid, md.m)))
            return false;
    } else {
        // Montgomery: no literal fallback, inputs must be canonical; slow
        // launches take the literal formulas whatever the values
        (void)big;
        __syncthreads();
    }
    TwView tv;
    tv.lds = nullptr;
    tv.gr31 = d0.twr;
    tv.glob = d0.tw;
    tv.W = 1;
    // the W = 2^wlog group prefixes of this item start at P0start (groups of
    // a multi-pass item are consecutive M, the first a multiple of G)
    const int wlog = item_tw_wlog(L, k, S, plog);
    const uint32_t P0start = (mlog == 0 || S <= plog) ? 0u : (uint32_t)(it.first & ((1ull << S) - 1ull));
    // stage the tables of as many first stages as the launch's area holds
    // (at least one radix-8 sub-pass' worth)
    int klds = 0;
    if (KIND != QK_SLOW)
        for (int j = k; j >= 3; j--)
            if (area.tw_cap >= tw_slots(j, wlog)) {
                klds = j;
                break;
            }
    if (klds > 0) {
        R31 *tws = reinterpret_cast<R31 *>(lds + area.tw_off);
        const uint64_t key = (uint64_t)S | ((uint64_t)klds << 8) | ((uint64_t)wlog << 16) | ((uint64_t)P0start << 32);
        if (s_twkey[0] != reinterpret_cast<uint64_t>(d0.twr) || s_twkey[1] != key) {
            __syncthreads();  // every thread has read the key
            stage_twiddles_r31(tws, d0.twr, S, klds, wlog, P0start);
            if (threadIdx.x == 0) {
                s_twkey[0] = reinterpret_cast<uint64_t>(d0.twr);
                s_twkey[1] = key;
            }
            __syncthreads();
        }
        tv.lds = tws;
        tv.W = 1u << wlog;
    }
    switch (elog) {
    case 1:
        run_network_elog<1, KIND>(lds, tv, k, S, mbase, klds, pmask, md);
        break;
    case 2:
        run_network_elog<2, KIND>(lds, tv, k, S, mbase, klds, pmask, md);
        break;
    default:
        run_network_elog<3, KIND>(lds, tv, k, S, mbase, klds, pmask, md);
        break;
    }
    // LDS -> HBM
    if (pair) {
        // position e of each transform's outputs: A's at tile position e, B's
        // at (H << k) + e; the product of the two when this is the last pass
        for (int j = 0; j < E / 2; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k;
            if (g >= nvalid) continue;
            const U128 ya = lds[lidx(e, k)];
            const U128 yb = lds[lidx((H << k) + e, k)];
            U128 *pa, *pb, *pc;
            if (mlog == 0) {
                const uint32_t ti = (uint32_t)(e & ((1 << k) - 1));
                pa = s_y[g] + ti;
                pb = s_y[H + g] + ti;
                pc = prod ? s_c[g] + ti : nullptr;
            } else {
                const uint32_t idx = ((uint32_t)it.first << k) + (uint32_t)e;
                pa = dst + idx;
                pb = dstB + idx;
                pc = prod ? d0.prod + idx : nullptr;
            }
            st16(pa, ya);
            st16(pb, yb);
            if (prod) st16(pc, pair_product<KIND>(ya, yb, md));
        }
        return true;
    }
    if (first_of_two(it)) {
        // the intermediate in the second pass's layout: consecutive threads
        // take consecutive groups at one position p
        const int G = 1 << Glog;
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e & (G - 1), pos = e >> Glog;
            if (g < nvalid) st16(dst + ((uint32_t)pos << mlog) + (uint32_t)it.first + (uint32_t)g, lds[lidx((g << k) + pos, k)]);
        }
        return true;
    }
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * kQueueThreads;
        if (contig) {
            if ((e >> k) < nvalid) st16(d0.y + e, lds[lidx(e, k)]);
        } else if (mlog == 0) {
            const int g = e >> k;
            if (g < nvalid) {
                const uint32_t ti = (uint32_t)(e & ((1 << k) - 1));
                if (scatter) {
                    // position ti of c takes network output brv(ti): contiguous
                    // stores, factor from the table permuted to positions
                    const U128 y = lds[lidx((g << k) + (int)bitrev(ti, (uint32_t)k), k)];
                    st16(s_y[g] + ti, montmul_r31<SPARSE>(y, ldr31(d0.scale + ((size_t)1 << k) + ti), md.mq));
                } else {
                    const U128 y = lds[lidx(e, k)];
                    st16(s_y[g] + ti, canon ? canonical(y, md.mq) : y);
                }
            }
        } else {
            if ((e >> k) < nvalid) {
                const uint32_t idx = ((uint32_t)it.first << k) + (uint32_t)e;
                if (scatter)
                    st16(dst + bitrev(idx, (uint32_t)L), montmul_r31<SPARSE>(lds[lidx(e, k)], ldr31(d0.scale + idx), md.mq));
                else if (canon)
                    st16(dst + idx, canonical(lds[lidx(e, k)], md.mq));
                else
                    st16(dst + idx, lds[lidx(e, k)]);
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Transform items of 2048-value tiles with the stage count K known at compile
// time (QK_FAST8 launches): 256 threads, 8 values each, G = 2^(11-K) groups of
// 2^K values, radix-8 register sub-passes (K mod 3 stages in a last radix-2 or
// radix-4 one), fast reference arithmetic, twiddle digits from device memory.
// With K fixed every LDS offset of a sub-pass is a constant distance from one
// per-thread base, and the tile loads and stores step by constants, so an item
// costs little beyond its butterflies.

// One register sub-pass: local stages [r0, r0 + KP) of the tile's groups.
template <int K, int KP, int KIND>
__device__ __forceinline__ void subpass8(U128 *lds, const R31 *__restrict__ twr, int S, int r0, uint32_t mbase,
                                         uint32_t pmask, int tid, const Mod &md)
{
    constexpr int Q = 1 << (3 - KP);  // sub-groups per thread
    constexpr int R = 1 << KP;
    constexpr int SL = K - KP;  // sub-groups per group = 2^SL
    U128 v[Q][R];
    R31 t[Q][R - 1];
    const uint32_t smask = (1u << S) - 1u;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = tid + q * kQueueThreads;
        const int g = sg >> SL;
        const int Mp = sg & ((1 << SL) - 1);
        // twiddles: stage r0 + rp, butterfly i uses tw[2^(S+r0+rp) + (P << rp) + (i mod 2^rp)]
        const uint32_t P = (((mbase + ((uint32_t)g & pmask)) & smask) << r0) | ((uint32_t)Mp & ((1u << r0) - 1u));
#pragma unroll
        for (int rp = 0; rp < KP; rp++) {
            const R31 *tb = twr + (1u << (S + r0 + rp)) + (P << rp);
#pragma unroll
            for (int j = 0; j < (1 << rp); j++) t[q][(1 << rp) - 1 + j] = ldr31(tb + j);
        }
        // the sub-group's values are 2^SL apart inside group g
        const int src = (g << K) + Mp;
#pragma unroll
        for (int j = 0; j < R; j++) v[q][j] = lds[lidx(src + (j << SL), K)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = tid + q * kQueueThreads;
        const int g = sg >> SL;
        const int Mp = sg & ((1 << SL) - 1);
        if constexpr (qk_mont(KIND))
            reg_network_mont<KP, qk_sparse(KIND)>(v[q], t[q], md);
        else
            reg_network_fast<KP>(v[q], t[q], md);
        // outputs of the sub-group: R consecutive values of one aligned block
        // of R (inside one aligned block of 8), so consecutive slots
        const int base = lidx((g << K) + (Mp << KP), K);
#pragma unroll
        for (int j = 0; j < R; j++) lds[base + j] = v[q][j];
    }
    __syncthreads();
}

template <int K, int KIND>
__device__ __forceinline__ bool do_ntt_item8(U128 *lds, const U128 **s_x, U128 **s_y, const U128 **s_x2, U128 **s_c,
                                             const WorkItem &it, const DevDesc *__restrict__ descs, const Mod &md)
{
    constexpr bool MONT = qk_mont(KIND), SPARSE = qk_sparse(KIND);
    constexpr int T = kQueueThreads;
    constexpr int Glog = 11 - K, G = 1 << Glog;
    constexpr int E = 8;
    // The thread index, opaque to the compiler here: every address below
    // derives from it, and as the plain index they would be loop-invariant
    // for the whole kernel, so the addresses of all nine stage counts would be
    // hoisted above the item loop and held in registers (spilling them).
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int L = it.L, S = it.S, mlog = L - K;
    const int nvalid = (int)it.nvalid;
    const DevDesc d0 = descs[it.desc];
    const bool pair = !MONT && (it.flags & IT_PAIR);
    const int plog = Glog - (pair ? 1 : 0);
    const uint32_t pmask = (1u << plog) - 1u;
    const int H = 1 << plog;
    const bool prod = pair && (it.flags & IT_PROD) && it.pass + 1 == it.np;
    const bool contig = !MONT && (it.flags & IT_CONTIG);
    // fused pointwise steps of the Montgomery path (a plain gather is a
    // product with B = 1)
    const bool gather = MONT && (it.flags & (IT_GATHER_MUL | IT_GATHER)) && it.pass == 0;
    const bool gmul = (it.flags & IT_GATHER_MUL) != 0;
    const bool scatter = MONT && (it.flags & IT_SCATTER_SCALE) && it.pass + 1 == it.np;
    const bool canon = MONT && (it.flags & IT_CANON) && it.pass + 1 == it.np;
    const U128 one{{1u, 0u, 0u, 0u}};
    const U128 *src, *srcB = nullptr;
    U128 *dst, *dstB = nullptr;
    pass_buffers(it, d0, src, dst);
    if (pair && mlog > 0) pass_buffers(it, descs[it.desc + 1], srcB, dstB);
    if (mlog == 0 && !contig) item_buffers<MONT>(s_x, s_y, s_x2, s_c, it, descs);

    // HBM -> LDS, every load of a thread issued before the first is used (an
    // element outside the kept groups reads a kept one and is dropped)
    U128 a[E], b[E];
    bool big = false;
    if (contig) {
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            a[j] = ld16(d0.x + ((e >> K) < nvalid ? e : 0));
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            if ((e >> K) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx(e, K)] = a[j];
            }
        }
    } else if (mlog == 0) {
        // whole transforms, one per group, each with its own buffers
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K, ti = e & ((1 << K) - 1);
            const int gv = (int)((uint32_t)g & pmask) < nvalid ? g : 0;
            a[j] = ld16(s_x[gv] + ti);
            if (MONT && gather && gmul) b[j] = ld16(s_x2[gv] + ti);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K, ti = e & ((1 << K) - 1);
            if ((int)((uint32_t)g & pmask) < nvalid) {
                if (MONT && gather) {
                    // network input brv(ti) is A[ti] * B[ti]: contiguous loads,
                    // bit-reversed placement
                    lds[lidx((g << K) + (int)bitrev((uint32_t)ti, K), K)] =
                        gather_product<SPARSE>(a[j], gmul ? b[j] : one, md.mq);
                } else {
                    big |= near_q(a[j], md.m);
                    lds[lidx(e, K)] = a[j];
                }
            }
        }
    } else if (MONT && gather) {
        // group g's inputs, bit-reversed, are the contiguous block
        // brv(M) << K of A and B (in the order brv(ti))
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K, jj = e & ((1 << K) - 1);
            const uint32_t gv = (uint32_t)(g < nvalid ? g : 0);
            const uint32_t blk = bitrev((uint32_t)it.first + gv, (uint32_t)mlog) << K;
            a[j] = ld16(d0.x + blk + jj);
            if (gmul) b[j] = ld16(d0.x2 + blk + jj);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K, jj = e & ((1 << K) - 1);
            if (g < nvalid)
                lds[lidx((g << K) + (int)bitrev((uint32_t)jj, K), K)] =
                    gather_product<SPARSE>(a[j], gmul ? b[j] : one, md.mq);
        }
    } else if (second_of_two(it)) {
        // contiguous groups (the first pass stored them so)
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K;
            const uint32_t gp = (uint32_t)g & pmask;
            const U128 *sb = (g >> plog) ? srcB : src;
            a[j] = ld16(sb + (((uint32_t)it.first + ((int)gp < nvalid ? gp : 0u)) << K) + (uint32_t)(e & ((1 << K) - 1)));
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            if ((int)((uint32_t)(e >> K) & pmask) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx(e, K)] = a[j];
            }
        }
    } else {
        // a thread keeps one group g (G divides the 256 threads) and loads
        // elements ti = (tid >> Glog) + j 2^(K-3) of it
        const int g = tid & (G - 1);
        const uint32_t gp = (uint32_t)g & pmask;
        const bool keep = (int)gp < nvalid;
        const U128 *sb = ((g >> plog) ? srcB : src) + (uint32_t)it.first + (keep ? gp : 0u);
        const int ti0 = tid >> Glog;
#pragma unroll
        for (int j = 0; j < E; j++) a[j] = ld16(sb + ((uint32_t)(ti0 + (j << (K - 3))) << mlog));
        if (keep) {
#pragma unroll
            for (int j = 0; j < E; j++) {
                big |= near_q(a[j], md.m);
                lds[lidx((g << K) + ti0 + (j << (K - 3)), K)] = a[j];
            }
        }
    }
    if constexpr (!MONT) {
        // a large twiddle, or a value above q: the literal formulas (the vote
        // is also the barrier after the tile writes)
        const bool twb = *(const __attribute__((address_space(1))) uint32_t *)d0.twflag != 0;
        if (__syncthreads_or(big || twb) && (twb || tile_above_q_kept(lds, K, Glog, pmask, nvalid, md.m)))
            return false;
    } else {
        // Montgomery: inputs are canonical, no literal path
        (void)big;
        __syncthreads();
    }

    const uint32_t mbase = (uint32_t)it.first;
#pragma unroll 1
    for (int r0 = 0; r0 + 3 <= K; r0 += 3) subpass8<K, 3, KIND>(lds, d0.twr, S, r0, mbase, pmask, tid, md);
    if constexpr (K % 3 != 0) subpass8<K, K % 3, KIND>(lds, d0.twr, S, K - K % 3, mbase, pmask, tid, md);

    // LDS -> HBM
    if (pair) {
#pragma unroll
        for (int j = 0; j < E / 2; j++) {
            const int e = tid + j * T;
            const int g = e >> K;
            if (g >= nvalid) continue;
            const U128 ya = lds[lidx(e, K)];
            const U128 yb = lds[lidx((H << K) + e, K)];
            U128 *pa, *pb, *pc;
            if (mlog == 0) {
                const uint32_t ti = (uint32_t)(e & ((1 << K) - 1));
                pa = s_y[g] + ti;
                pb = s_y[H + g] + ti;
                pc = s_c[g] + ti;
            } else {
                const uint32_t idx = ((uint32_t)it.first << K) + (uint32_t)e;
                pa = dst + idx;
                pb = dstB + idx;
                pc = d0.prod + idx;
            }
            st16(pa, ya);
            st16(pb, yb);
            if (prod) st16(pc, mulmod_r31(yb, to_r31(ya), md.c));
        }
        return true;
    }
    if (first_of_two(it)) {
        // the intermediate in the second pass's layout: a thread keeps group
        // g and stores its positions (tid >> Glog) + j 2^(K-3)
        const int g = tid & (G - 1);
        const int p0 = tid >> Glog;
        if (g < nvalid) {
            U128 *db = dst + (uint32_t)it.first + (uint32_t)g;
#pragma unroll
            for (int j = 0; j < E; j++) {
                const int pos = p0 + (j << (K - 3));
                st16(db + ((uint32_t)pos << mlog), lds[lidx((g << K) + pos, K)]);
            }
        }
        return true;
    }
    if (contig || mlog > 0) {
        // the kept groups' outputs are one contiguous range (in network
        // order; a Montgomery scatter stores output idx at brv(idx))
        const uint32_t first = (uint32_t)it.first << K;
        U128 *yb = contig ? d0.y : dst + first;
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            if ((e >> K) < nvalid) {
                const U128 v = lds[lidx(e, K)];
                if (MONT && scatter)
                    st16(dst + bitrev(first + (uint32_t)e, (uint32_t)L),
                         montmul_r31<SPARSE>(v, ldr31(d0.scale + first + (uint32_t)e), md.mq));
                else if (MONT && canon)
                    st16(yb + e, canonical(v, md.mq));
                else
                    st16(yb + e, v);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = tid + j * T;
            const int g = e >> K;
            if (g < nvalid) {
                const uint32_t ti = (uint32_t)(e & ((1 << K) - 1));
                if (MONT && scatter) {
                    // position ti of c takes network output brv(ti): contiguous
                    // stores, factor from the table permuted to positions
                    const U128 v = lds[lidx((g << K) + (int)bitrev(ti, K), K)];
                    st16(s_y[g] + ti, montmul_r31<SPARSE>(v, ldr31(d0.scale + (1u << K) + ti), md.mq));
                } else {
                    const U128 v = lds[lidx(e, K)];
                    st16(s_y[g] + ti, (MONT && canon) ? canonical(v, md.mq) : v);
                }
            }
        }
    }
    return true;
}

// A transform item of a launch of kind qk_item8: the code for its stage count.
template <int KIND>
__device__ __forceinline__ bool do_ntt_item8_k(U128 *lds, const U128 **s_x, U128 **s_y, const U128 **s_x2,
                                               U128 **s_c, const WorkItem &it, const DevDesc *__restrict__ descs,
                                               const Mod &md)
{
    switch (it.k) {
    case 3:
        return do_ntt_item8<3, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 4:
        return do_ntt_item8<4, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 5:
        return do_ntt_item8<5, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 6:
        return do_ntt_item8<6, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 7:
        return do_ntt_item8<7, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 8:
        return do_ntt_item8<8, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 9:
        return do_ntt_item8<9, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    case 10:
        return do_ntt_item8<10, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    default:
        return do_ntt_item8<11, KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
    }
}

// One product item (reference arithmetic).  A fast launch leaves an item with
// an operand >= 2^124 to the literal formulas (returns false, nothing written).
template <int KIND>
__device__ __forceinline__ bool do_mul_item(const WorkItem &it, const DevMul *__restrict__ muls, const Mod &md)
{
    constexpr int J = kQueueTile / kQueueThreads;
    const DevMul mu = muls[it.desc];
    if constexpr (qk_fast(KIND)) {
        U128 a[J], b[J];
        bool big = false;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t e = threadIdx.x + j * kQueueThreads;
            if (e < it.nvalid) {
                a[j] = ld16(mu.a + it.first + e);
                b[j] = ld16(mu.b + it.first + e);
                big |= !below_2_124(a[j], b[j]);
            }
        }
        if (__syncthreads_or(big)) return false;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t e = threadIdx.x + j * kQueueThreads;
            if (e < it.nvalid) st16(mu.c + it.first + e, mulmod_r31(b[j], to_r31(a[j]), md.c));
        }
    } else {
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t e = threadIdx.x + j * kQueueThreads;
            if (e < it.nvalid) {
                const size_t i = it.first + e;
                st16(mu.c + i, mulmod_ref(ld16(mu.a + i), ld16(mu.b + i), md.m, md.mu));
            }
        }
    }
    return true;
}

template <int KIND>
__global__ __launch_bounds__(kQueueThreads, 3) void queue_kernel(const WorkItem *__restrict__ items, uint32_t nitems,
                                                                 const DevDesc *__restrict__ descs,
                                                                 const DevMul *__restrict__ muls, QueueCtl *ctl,
                                                                 TwArea area, Mod md, PuntList punt)
{
    constexpr bool MONT = qk_mont(KIND);
    extern __shared__ U128 lds[];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_twkey[2];  // tables staged in the twiddle area (none yet)
    __shared__ const U128 *s_x[kQueueThreads];  // buffers of the transforms of a packed item
    __shared__ U128 *s_y[kQueueThreads];
    __shared__ const U128 *s_x2[MONT ? kQueueThreads : 1];
    __shared__ U128 *s_c[MONT ? 1 : kQueueThreads / 2];  // products of a paired item
    if (threadIdx.x == 0) s_twkey[0] = s_twkey[1] = ~0ull;
    // A slow launch over a punt list takes the items the fast launch before
    // it on the stream appended (that launch has completed).  The workgroup
    // that reads the count last clears it for the next execution.
    if constexpr (KIND == QK_SLOW)
        if (punt.idx) {
            // one read per workgroup, shared with its threads, so that they
            // all see the same count
            // (a zero count needs no clearing: then no workgroup counts itself;
            // a nonzero one is read as such by every workgroup, since it is
            // cleared only after all have read it)
            if (threadIdx.x == 0) {
                s_ticket = __hip_atomic_load(punt.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (s_ticket != 0 &&
                    __hip_atomic_fetch_add(&ctl->readers, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                        gridDim.x - 1) {
                    __hip_atomic_store(punt.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&ctl->readers, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
            nitems = s_ticket;
            __syncthreads();
        }
    // Workgroup b starts with item b; after each item it draws from the
    // counter, draw c giving item gridDim + c (unless the grid has a
    // workgroup per item).  A device-wide atomic is
    // performed beyond the XCDs' L2 caches, so the first round draws nothing.
    // Every processed item is followed by exactly one draw, so the draws are
    // 0 .. nitems-1: the workgroup drawing nitems-1 is the last to touch the
    // counter and rearms it for the next launch on this counter block.
    uint32_t t = blockIdx.x;
    while (t < nitems) {
        const WorkItem it = items[(KIND == QK_SLOW && punt.idx) ? punt.idx[t] : t];
        bool done = true;
        if (it.kind == ITEM_NTT) {
            if constexpr (qk_item8(KIND))
                done = do_ntt_item8_k<KIND>(lds, s_x, s_y, s_x2, s_c, it, descs, md);
            else
                done = do_ntt_item<KIND>(lds, s_x, s_y, s_x2, s_c, s_twkey, it, descs, area, md);
        } else if constexpr (!MONT) {
            done = do_mul_item<KIND>(it, muls, md);
        }
        if (threadIdx.x == 0 && gridDim.x >= nitems) {
            // one workgroup per item (a level of equal items): no draws
            if constexpr (qk_fast(KIND))
                if (!done) punt.idx[atomicAdd(punt.count, 1u)] = t;
            s_ticket = nitems;
        } else if (threadIdx.x == 0) {
            if constexpr (qk_fast(KIND))
                if (!done) punt.idx[atomicAdd(punt.count, 1u)] = t;
            const uint32_t c = atomicAdd(&ctl->ticket, 1u);
            if (c == nitems - 1) atomicExch(&ctl->ticket, 0u);
            s_ticket = gridDim.x + c;
        }
        __syncthreads();
        t = s_ticket;
        __syncthreads();
    }
}

}  // namespace bntt
