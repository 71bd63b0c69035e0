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
//   QK_MONT  exact Montgomery arithmetic (polynomial-multiply server);
//            QK_MONT_S the same for a modulus whose 31-bit digits 1 and 2
//            are zero (a shorter reduction)
// Keeping the literal formulas out of the fast kernel keeps its registers
// free of spills.  Kinds of item:
//   ITEM_NTT   stages [S, S+k) of the forward network for G groups (several
//              whole transforms when k = log2 N, else blocks of one transform)
//   ITEM_MUL   2048 elements of c = a * b mod q
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
enum ItemFlags : uint8_t { IT_GATHER_MUL = 1, IT_SCATTER_SCALE = 2, IT_CONTIG = 4, IT_GATHER = 8, IT_CANON = 16 };

struct DevDesc {
    const U128 *x;
    U128 *y;
    U128 *tmp;
    const U128 *tw;
    const R31 *twr;          // digit copy of tw (reference arithmetic) or tw*R digits (Montgomery)
    const uint32_t *twflag;  // nonzero when some entry of tw (past entry 0) is >= 2^124
    const U128 *x2;          // IT_GATHER_MUL: second factor B (x holds A)
    const R31 *scale;        // IT_SCATTER_SCALE: S in digits
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
    uint32_t ticket;  // dynamic draws of the running launch; 0 between launches
    uint32_t pad[3];
};

enum QueueKind : int { QK_FAST = 0, QK_SLOW = 1, QK_MONT = 2, QK_MONT_S = 3 };
__host__ __device__ constexpr bool qk_mont(int kind) { return kind == QK_MONT || kind == QK_MONT_S; }

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
                                                 const Mod &md)
{
    if constexpr (qk_mont(KIND)) {
        constexpr int M = KIND == QK_MONT_S ? TW_MONT_S : TW_MONT;
        constexpr int ML = KIND == QK_MONT_S ? TW_MONT_S_LDS : TW_MONT_LDS;
        if (klds > 0)
            run_group_network<ELOG, ML, true, M>(lds, tv, k, S, mbase, md, klds);
        else
            run_group_network<ELOG, M, true>(lds, tv, k, S, mbase, md);
    } else if constexpr (KIND == QK_SLOW) {
        // the literal Barrett product whatever the modulus
        run_group_network<ELOG, TW_SLOW, false>(lds, tv, k, S, mbase, md);
    } else {
        if (klds > 0)
            run_group_network<ELOG, TW_LDS, true, TW_GLOBAL>(lds, tv, k, S, mbase, md, klds);
        else
            run_group_network<ELOG, TW_GLOBAL, true>(lds, tv, k, S, mbase, md);
    }
}

// LDS layout of a launch: the data tile in U128 slots [0, tw_off), then room
// for tw_cap staged digit twiddles.  The staged tables stay valid across the
// items of a workgroup until an item needs different ones (key).
struct TwArea {
    uint32_t tw_off;
    uint32_t tw_cap;
};

// Twiddle-table prefixes staged for a transform item (see do_ntt_item).
__host__ __device__ inline int item_tw_wlog(int L, int k, int S, int Glog)
{
    return L == k ? 0 : (Glog < S ? Glog : S);
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
                                          const WorkItem &it, const DevDesc &d0, const U128 *src, const Mod &md)
{
    constexpr bool MONT = qk_mont(KIND), SPARSE = KIND == QK_MONT_S;
    const int L = it.L, k = it.k, Glog = it.Glog, mlog = L - k, nvalid = (int)it.nvalid;
    const int G = 1 << Glog;
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
            const int gv = g < nvalid ? g : 0;
            a[j] = ld16(s_x[gv] + ti);
            if (gather && gmul) b[j] = ld16(s_x2[gv] + ti);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e >> k, ti = e & ((1 << k) - 1);
            if (g < nvalid) {
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
    } else {
        // blocks of one transform: element ti of group g at (ti << mlog) + M
        // (indices within one transform, at most 2^24 points, fit 32 bits)
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e & (G - 1), ti = e >> Glog;
            const uint32_t gv = (uint32_t)(g < nvalid ? g : 0);
            a[j] = ld16(src + ((uint32_t)ti << mlog) + (uint32_t)it.first + gv);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * kQueueThreads;
            const int g = e & (G - 1), ti = e >> Glog;
            if (g < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx((g << k) + ti, k)] = a[j];
            }
        }
    }
    return big;
}

// One transform item: load, network, store.  Returns false when a fast
// launch leaves the item to the literal formulas (nothing written).
template <int KIND>
__device__ bool do_ntt_item(U128 *lds, const U128 **s_x, U128 **s_y, const U128 **s_x2, uint64_t *s_twkey,
                            const WorkItem &it, const DevDesc *__restrict__ descs, TwArea area, const Mod &md)
{
    constexpr bool MONT = qk_mont(KIND), SPARSE = KIND == QK_MONT_S;
    const int L = it.L, S = it.S, k = it.k, Glog = it.Glog, elog = it.elog;
    const int G = 1 << Glog, E = 1 << elog;
    const int mlog = L - k;
    const int nvalid = (int)it.nvalid;
    bool big = false;
    const DevDesc d0 = descs[it.desc];
    const U128 *src;
    U128 *dst;
    if (it.np == 1) {
        src = d0.x;
        dst = d0.y;
    } else {
        const int p = it.pass, np = it.np;
        dst = ((np - 1 - p) % 2 == 0) ? d0.y : d0.tmp;
        src = (p == 0) ? d0.x : (((np - p) % 2 == 0) ? d0.y : d0.tmp);
    }
    // fused pointwise steps (Montgomery products only); a plain gather is a
    // product with B = 1
    const bool gather = MONT && (it.flags & (IT_GATHER_MUL | IT_GATHER)) && it.pass == 0;
    const bool gmul = (it.flags & IT_GATHER_MUL) != 0;
    const U128 one{{1u, 0u, 0u, 0u}};
    const bool scatter = MONT && (it.flags & IT_SCATTER_SCALE) && it.pass + 1 == it.np;
    const bool canon = MONT && (it.flags & IT_CANON) && it.pass + 1 == it.np;
    const bool contig = !MONT && (it.flags & IT_CONTIG);
    if (mlog == 0 && !contig) {
        // whole transforms, one per group, each with its own buffers
        if ((int)threadIdx.x < nvalid) {
            const DevDesc &d = descs[it.desc + threadIdx.x];
            s_x[threadIdx.x] = d.x;
            s_y[threadIdx.x] = d.y;
            if constexpr (MONT) s_x2[threadIdx.x] = d.x2;
        }
        __syncthreads();
    }
    // HBM -> LDS
    switch (elog) {
    case 1:
        big = load_tile<2, KIND>(lds, s_x, s_x2, it, d0, src, md);
        break;
    case 2:
        big = load_tile<4, KIND>(lds, s_x, s_x2, it, d0, src, md);
        break;
    default:
        big = load_tile<8, KIND>(lds, s_x, s_x2, it, d0, src, md);
        break;
    }
    const uint32_t mbase = (uint32_t)it.first;
    if constexpr (KIND == QK_FAST) {
        // a large twiddle, or a value above q (decided exactly in the rare
        // workgroups where some value is near q): the literal formulas.  The
        // vote is also the barrier after the tile writes.
        const bool twb = *(const __attribute__((address_space(1))) uint32_t *)d0.twflag != 0;  // one flag per item
        if (__syncthreads_or(big || twb) && (twb || tile_above_q(lds, nvalid << k, k, md.m))) return false;
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
    const int wlog = (mlog == 0) ? 0 : (Glog < S ? Glog : S);
    const uint32_t P0start = (mlog == 0 || S <= Glog) ? 0u : (uint32_t)(it.first & ((1ull << S) - 1ull));
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
        run_network_elog<1, KIND>(lds, tv, k, S, mbase, klds, md);
        break;
    case 2:
        run_network_elog<2, KIND>(lds, tv, k, S, mbase, klds, md);
        break;
    default:
        run_network_elog<3, KIND>(lds, tv, k, S, mbase, klds, md);
        break;
    }
    // LDS -> HBM
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

// One product item (reference arithmetic).  A fast launch leaves an item with
// an operand >= 2^124 to the literal formulas (returns false, nothing written).
template <int KIND>
__device__ __forceinline__ bool do_mul_item(const WorkItem &it, const DevMul *__restrict__ muls, const Mod &md)
{
    constexpr int J = kQueueTile / kQueueThreads;
    const DevMul mu = muls[it.desc];
    if constexpr (KIND == QK_FAST) {
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
    if (threadIdx.x == 0) s_twkey[0] = s_twkey[1] = ~0ull;
    // a slow launch over a punt list takes the items the fast launch before
    // it on the stream appended (that launch has completed)
    if constexpr (KIND == QK_SLOW)
        if (punt.idx) nitems = __hip_atomic_load(punt.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Workgroup b starts with item b; after each item it draws from the
    // counter, draw c giving item gridDim + c.  A device-wide atomic is
    // performed beyond the XCDs' L2 caches, so the first round draws nothing.
    // Every processed item is followed by exactly one draw, so the draws are
    // 0 .. nitems-1: the workgroup drawing nitems-1 is the last to touch the
    // counter and rearms it for the next launch on this counter block.
    uint32_t t = blockIdx.x;
    while (t < nitems) {
        const WorkItem it = items[(KIND == QK_SLOW && punt.idx) ? punt.idx[t] : t];
        bool done = true;
        if (it.kind == ITEM_NTT) {
            done = do_ntt_item<KIND>(lds, s_x, s_y, s_x2, s_twkey, it, descs, area, md);
        } else if constexpr (!MONT) {
            done = do_mul_item<KIND>(it, muls, md);
        }
        if (threadIdx.x == 0) {
            if constexpr (KIND == QK_FAST)
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
