// queue.hpp -- the persistent work-queue kernel and its work items, shared by
// the uneven-batch planner (batch.hip, reference arithmetic) and the
// polynomial-multiply server (serve.hip, exact Montgomery arithmetic).
//
// A launch processes one list of equal-size work items: a grid sized to the
// resident capacity of the chip whose workgroups take items from an atomic
// ticket until the list is exhausted; the last workgroup to leave rearms the
// counters for the next launch.  Kinds of item:
//   ITEM_NTT   stages [S, S+k) of the forward network for G groups (several
//              whole transforms when k = log2 N, else blocks of one transform)
//   ITEM_MUL   2048 elements of c = a * b mod q
//   ITEM_GMUL  2048 elements of G[i] = A[brv(i)] * B[brv(i)] (Montgomery)
//   ITEM_SCAT  2048 elements of C[brv(i)] = Y[i] * S[i] (Montgomery)
// where brv reverses the log2 N index bits of a transform.
#pragma once

#include "ntt_core.hpp"

namespace bntt {

constexpr int kQueueThreads = 256;
constexpr int kQueueMaxElog = 3;
constexpr int kQueueTile = kQueueThreads << kQueueMaxElog;  // 2048 values
// padded data tile (the twiddles are read from device memory)
constexpr size_t kQueueMaxLds = 64 * 1024;

enum ItemKind : uint8_t { ITEM_NTT = 0, ITEM_MUL = 1, ITEM_GMUL = 2, ITEM_SCAT = 3 };

struct DevDesc {
    const U128 *x;
    U128 *y;
    U128 *tmp;
    const U128 *tw;
    const R31 *twr;          // digit copy of tw (reference arithmetic) or tw*R digits (Montgomery)
    const uint32_t *twflag;  // nonzero when some entry of tw (past entry 0) is not small
};

struct DevMul {
    const U128 *a;
    const U128 *b;
    U128 *c;
    uint64_t n;
};

// Operands of a permuted pointwise product (ITEM_GMUL / ITEM_SCAT).
struct DevPoint {
    const U128 *src0;
    const U128 *src1;  // GMUL: second operand
    const R31 *tab;    // SCAT: per-index factor in digits
    U128 *dst;
    uint32_t log2n;
    uint32_t pad;
};

struct WorkItem {
    uint8_t kind, L, S, k;
    uint8_t Glog, elog, pass, np;
    uint32_t desc;    // descriptor / product / point index
    uint32_t nvalid;  // groups (NTT) or elements (products)
    uint64_t first;   // first group of the pass (NTT multi-pass) or element offset (products)
    uint64_t pad;
};
static_assert(sizeof(WorkItem) == 32, "work item layout");

struct QueueCtl {
    uint32_t ticket;
    uint32_t done;
    uint32_t pad[2];
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

template <int ELOG, bool MONT>
__device__ __forceinline__ void run_network_elog(U128 *lds, const TwView &tv, int k, int S, uint32_t mbase, bool slow,
                                                 const Mod &md)
{
    if constexpr (MONT) {
        run_group_network<ELOG, TW_MONT, true>(lds, tv, k, S, mbase, md);
    } else {
        if (!slow)
            run_group_network<ELOG, TW_GLOBAL, true>(lds, tv, k, S, mbase, md);
        else
            run_group_network<ELOG, TW_SLOW, true>(lds, tv, k, S, mbase, md);
    }
}

// One transform item: load, network, store.
template <bool MONT>
__device__ void do_ntt_item(U128 *lds, const U128 **s_x, U128 **s_y, const WorkItem &it,
                            const DevDesc *__restrict__ descs, const Mod &md)
{
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
    if (mlog == 0) {
        // whole transforms, one per group, each with its own buffers
        if ((int)threadIdx.x < nvalid) {
            s_x[threadIdx.x] = descs[it.desc + threadIdx.x].x;
            s_y[threadIdx.x] = descs[it.desc + threadIdx.x].y;
        }
        __syncthreads();
    }
    // HBM -> LDS
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * kQueueThreads;
        if (mlog == 0) {
            const int g = e >> k, ti = e & ((1 << k) - 1);
            if (g < nvalid) {
                const U128 a = ld16(s_x[g] + ti);
                big |= is_big(a);
                lds[lidx(e, k)] = a;
            }
        } else {
            const int g = e & (G - 1), ti = e >> Glog;
            if (g < nvalid) {
                const size_t M = it.first + (size_t)g;
                const U128 a = ld16(src + ((size_t)ti << mlog) + M);
                big |= is_big(a);
                lds[lidx((g << k) + ti, k)] = a;
            }
        }
    }
    const uint32_t mbase = (uint32_t)it.first;
    bool slow = false;
    if constexpr (MONT) {
        // the Montgomery path has no literal fallback: inputs must be canonical
        (void)big;
        __syncthreads();
    } else {
        big |= *d0.twflag != 0;
        slow = __syncthreads_or(big) || md.pm == 0;
    }
    TwView tv;
    tv.lds = nullptr;
    tv.gr31 = d0.twr;
    tv.glob = d0.tw;
    tv.W = 1;
    switch (elog) {
    case 1:
        run_network_elog<1, MONT>(lds, tv, k, S, mbase, slow, md);
        break;
    case 2:
        run_network_elog<2, MONT>(lds, tv, k, S, mbase, slow, md);
        break;
    default:
        run_network_elog<3, MONT>(lds, tv, k, S, mbase, slow, md);
        break;
    }
    // LDS -> HBM
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * kQueueThreads;
        if (mlog == 0) {
            const int g = e >> k;
            if (g < nvalid) st16(s_y[g] + (e & ((1 << k) - 1)), lds[lidx(e, k)]);
        } else {
            if ((e >> k) < nvalid) st16(dst + (it.first << k) + e, lds[lidx(e, k)]);
        }
    }
}

__device__ __forceinline__ void do_mul_item(const WorkItem &it, const DevMul *__restrict__ muls, const Mod &md)
{
    const DevMul mu = muls[it.desc];
#pragma unroll
    for (int j = 0; j < kQueueTile / kQueueThreads; j++) {
        const uint32_t e = threadIdx.x + j * kQueueThreads;
        if (e < it.nvalid) {
            const size_t i = it.first + e;
            const U128 a = ld16(mu.a + i), b = ld16(mu.b + i);
            st16(mu.c + i, md.pm ? mulmod_elem<true>(a, b, md) : mulmod_elem<false>(a, b, md));
        }
    }
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, uint32_t log2n)
{
    return log2n ? (__brev(i) >> (32u - log2n)) : 0u;
}

// G[i] = A[brv(i)] * B[brv(i)] * R^-1  or  C[brv(i)] = Y[i] * S[i] * R^-1.
template <int KIND>
__device__ __forceinline__ void do_point_item(const WorkItem &it, const DevPoint *__restrict__ pts, const Mod &md)
{
    const DevPoint pt = pts[it.desc];
#pragma unroll
    for (int j = 0; j < kQueueTile / kQueueThreads; j++) {
        const uint32_t e = threadIdx.x + j * kQueueThreads;
        if (e < it.nvalid) {
            const uint32_t i = (uint32_t)it.first + e;
            const uint32_t r = bitrev(i, pt.log2n);
            if constexpr (KIND == ITEM_GMUL) {
                const U128 a = ld16(pt.src0 + r), b = ld16(pt.src1 + r);
                st16(pt.dst + i, montmul_r31(a, to_r31(b), md.mq));
            } else {
                const U128 y = ld16(pt.src0 + i);
                st16(pt.dst + r, montmul_r31(y, pt.tab[i], md.mq));
            }
        }
    }
}

template <bool MONT>
__global__ __launch_bounds__(kQueueThreads, 3) void queue_kernel(const WorkItem *__restrict__ items, uint32_t nitems,
                                                                 const DevDesc *__restrict__ descs,
                                                                 const DevMul *__restrict__ muls,
                                                                 const DevPoint *__restrict__ pts, QueueCtl *ctl,
                                                                 Mod md)
{
    extern __shared__ U128 lds[];
    __shared__ uint32_t s_ticket;
    __shared__ const U128 *s_x[kQueueThreads];  // buffers of the transforms of a packed item
    __shared__ U128 *s_y[kQueueThreads];
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(&ctl->ticket, 1u);
        __syncthreads();
        const uint32_t t = s_ticket;
        __syncthreads();
        if (t >= nitems) break;
        const WorkItem it = items[t];
        if (it.kind == ITEM_NTT) {
            do_ntt_item<MONT>(lds, s_x, s_y, it, descs, md);
        } else if constexpr (MONT) {
            if (it.kind == ITEM_GMUL)
                do_point_item<ITEM_GMUL>(it, pts, md);
            else
                do_point_item<ITEM_SCAT>(it, pts, md);
        } else {
            do_mul_item(it, muls, md);
        }
    }
    // every workgroup has drawn a ticket past the end; the last one to leave
    // rearms the counters for the next launch on this counter block
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t d = atomicAdd(&ctl->done, 1u);
        if (d == gridDim.x - 1) {
            atomicExch(&ctl->ticket, 0u);
            atomicExch(&ctl->done, 0u);
        }
    }
}

}  // namespace bntt
