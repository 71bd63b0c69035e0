// ntt.hip -- launch-per-call forward NTT: `count` transforms of one size with
// one twiddle table, in one launch (N <= 4096) or two or three launches
// through device memory (N > 4096, and a lone transform of 2^11 or 2^12
// points given a tmp).  The network and its blocking are described in
// ntt_core.hpp.
#include "internal.hpp"
#include "ntt_core.hpp"

namespace bntt {

// One pass: stages [S, S+k) of `count` transforms of size 2^L.
// Group id gamma = transform * 2^(L-k) + M; this workgroup owns groups
// [blockIdx.x * G, blockIdx.x * G + G).  Block size = (G << k) >> ELOG.
template <int ELOG, bool PM>
__global__ __launch_bounds__(kMaxThreads) void ntt_pass_kernel(U128 *__restrict__ y, const U128 *__restrict__ x,
                                                               const U128 *__restrict__ tw, int L, int S, int k,
                                                               int Glog, int klds, size_t total_groups, Mod md,
                                                               uint32_t *done, uint32_t seq)
{
    extern __shared__ U128 lds[];
    constexpr int E = 1 << ELOG;
    const int T = blockDim.x;
    const int G = 1 << Glog;
    const int mlog = L - k;  // groups per transform = 2^mlog
    const size_t gamma0 = (size_t)blockIdx.x << Glog;
    const int nvalid = (int)((total_groups - gamma0) < (size_t)G ? (total_groups - gamma0) : (size_t)G);
    bool big = false;

    // HBM -> LDS.  Single-pass (k == L): groups are whole transforms and the
    // block's data is one contiguous range.  Otherwise element T_idx of group g
    // sits at transform*N + T_idx*2^mlog + M; consecutive threads take
    // consecutive g (consecutive addresses).  All E loads of a thread are
    // issued before the first is used (an element past the block's groups
    // reads the block's first group instead, and is not kept), so they are in
    // flight together rather than one at a time.
    U128 a[E];
    if (mlog == 0) {
        // the workgroup's G transforms are one contiguous range
        const U128 *xb = x + (gamma0 << L);
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * T;
            a[j] = ld16(xb + ((e >> k) < nvalid ? e : 0));
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * T;
            if ((e >> k) < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx(e, k)] = a[j];
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * T;
            const int g = e & (G - 1), ti = e >> Glog;
            const size_t gamma = gamma0 + (size_t)(g < nvalid ? g : 0);
            const size_t tr = gamma >> mlog;
            const size_t M = gamma & (((size_t)1 << mlog) - 1);
            a[j] = ld16(x + (tr << L) + ((size_t)ti << mlog) + M);
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * T;
            const int g = e & (G - 1), ti = e >> Glog;
            if (g < nvalid) {
                big |= near_q(a[j], md.m);
                lds[lidx((g << k) + ti, k)] = a[j];
            }
        }
    }

    // Stage the twiddles this workgroup uses, as digits.
    const uint32_t mbase = (uint32_t)(gamma0 & (((size_t)1 << mlog) - 1));
    const int wlog = Glog < S ? Glog : S;
    const uint32_t W = 1u << wlog;
    const uint32_t P0start = (wlog == Glog) ? (mbase & ((1u << S) - 1u)) : 0u;
    R31 *tws = reinterpret_cast<R31 *>(lds + lds_data_slots(k, Glog));
    const bool twbig = stage_twiddles(tws, tw, S, k, wlog, P0start, klds);
    // the vote is also the barrier after the tile and table writes; a value
    // near q (rare for residues) is decided on the exact values
    bool slow = __syncthreads_or(big || twbig);
    if (slow && PM) slow = __syncthreads_or(twbig) || tile_above_q(lds, nvalid << k, k, md.m);
    slow = slow || !PM;

    TwView tv;
    tv.lds = tws;
    tv.gr31 = nullptr;
    tv.glob = tw;
    tv.W = W;
    if (!slow) {
        if constexpr (PM) run_group_network<ELOG, TW_LDS, PM, TW_G128>(lds, tv, k, S, mbase, md, klds);
    } else {
        run_group_network<ELOG, TW_SLOW, PM>(lds, tv, k, S, mbase, md);
    }

    // LDS -> HBM: the G groups' outputs are one contiguous range.
    U128 *yb = y + (gamma0 << k);
#pragma unroll
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * T;
        if ((e >> k) < nvalid) st16(yb + e, lds[lidx(e, k)]);
    }
    if (done) {
        // a one-workgroup launch whose completion the host reads from pinned
        // memory: every wave makes its stores visible system-wide, then one
        // release store posts the sequence number
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <int ELOG, bool PM>
static hipError_t launch_pass(U128 *y, const U128 *x, const U128 *tw, int L, int S, int k, int Glog, size_t count,
                              const Mod &md, hipStream_t stream, uint32_t *done = nullptr, uint32_t seq = 0)
{
    const size_t total_groups = count << (L - k);
    const size_t blocks = (total_groups + ((size_t)1 << Glog) - 1) >> Glog;
    const int threads = (1 << (Glog + k)) >> ELOG;
    // stage the tables of as many stages as keep three 256-thread workgroups
    // on a CU; the remaining stages read the caller's table
    int klds = k;
    if (threads == 256)
        while (klds > 1 && pass_lds_bytes(k, S, Glog, klds) > kLds3Bytes) klds--;
    const size_t lds_bytes = pass_lds_bytes(k, S, Glog, klds);
    if (threads < 1 || threads > kMaxThreads || blocks > 0x7fffffffu || lds_bytes > kMaxLdsBytes)
        return hipErrorInvalidConfiguration;
    if (done && blocks != 1) return hipErrorInvalidValue;
    // raise the dynamic LDS limit once per instantiation and device
    static std::atomic<uint64_t> lds_done{0};
    const hipError_t attr =
        bntt_allow_lds(reinterpret_cast<const void *>(&ntt_pass_kernel<ELOG, PM>), (int)kMaxLdsBytes, lds_done);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL((ntt_pass_kernel<ELOG, PM>), dim3((unsigned)blocks), dim3(threads), lds_bytes, stream, y, x,
                       tw, L, S, k, Glog, klds, total_groups, md, done, seq);
    return hipGetLastError();
}

template <bool PM>
static hipError_t launch_pass_e(U128 *y, const U128 *x, const U128 *tw, int L, int S, int k, int Glog, int elog,
                                size_t count, const Mod &md, hipStream_t stream, uint32_t *done = nullptr,
                                uint32_t seq = 0)
{
    switch (elog) {
    case 1:
        return launch_pass<1, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq);
    case 2:
        return launch_pass<2, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq);
    default:
        return launch_pass<3, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq);
    }
}

// Transforms per launch below which the single-pass kernel gives each
// transform its own workgroup of up to 512 threads (latency-bound regime)
// instead of packing 2048 values of several transforms per workgroup.
constexpr size_t kFewTransforms = 64;

hipError_t launch_ntt_uniform(U128 *y, const U128 *x, const U128 *tw, int L, size_t count, const Mod &md, U128 *tmp,
                              hipStream_t stream, uint32_t *done, uint32_t seq)
{
    if (count == 0) return hipSuccess;
    if (L < kMinLog2 || L > kMaxLog2) return hipErrorInvalidValue;
    // a completion flag needs a launch of one workgroup: one transform of at
    // most 4096 points
    if (done && (count != 1 || L > kLdsMaxLog2)) return hipErrorInvalidValue;
    const bool pm = md.pm != 0;
    // A lone transform of 2^11 or 2^12 points on one workgroup is
    // latency-bound (15 and 26 us); given a tmp, it takes two passes over
    // several CUs instead.
    const bool spread = tmp != nullptr && count == 1 && L >= kSpreadMinLog2 && done == nullptr;
    if (L <= kLdsMaxLog2 && !spread) {
        int elog, Glog;
        if (count <= kFewTransforms) {
            // one transform per workgroup, as many threads as the network allows
            elog = L > 10 ? L - 9 : 1;
            Glog = 0;
        } else {
            // 256 threads of up to 8 points, several transforms per workgroup
            elog = L < 3 ? L : 3;
            Glog = 8 + elog - L;
            if (Glog < 0) Glog = 0;
        }
        return pm ? launch_pass_e<true>(y, x, tw, L, 0, L, Glog, elog, count, md, stream, done, seq)
                  : launch_pass_e<false>(y, x, tw, L, 0, L, Glog, elog, count, md, stream, done, seq);
    }
    if (tmp == nullptr) return hipErrorInvalidValue;
    const int np = L <= 16 ? 2 : 3;
    // 2048 values per workgroup (radix-8 sub-passes) when a pass has at least
    // 2048 such tiles; fewer values per workgroup below that, so that the
    // passes of a lone transform spread over more CUs
    const size_t total = count << L;
    const int elog = total >= ((size_t)1 << 22) ? 3 : total >= ((size_t)1 << 21) ? 2 : 1;
    const int tile_log = 8 + elog;
    int ks[3];
    if (!pass_split(L, np, elog, tile_log, ks)) return hipErrorInvalidValue;
    int S = 0;
    for (int p = 0; p < np; p++) {
        const int k = ks[p];
        const int Glog = tile_log - k;
        const U128 *src = (p == 0) ? x : ((np - p) % 2 == 0 ? y : tmp);
        U128 *dst = ((np - 1 - p) % 2 == 0) ? y : tmp;
        const hipError_t e = pm ? launch_pass_e<true>(dst, src, tw, L, S, k, Glog, elog, count, md, stream)
                                : launch_pass_e<false>(dst, src, tw, L, S, k, Glog, elog, count, md, stream);
        if (e != hipSuccess) return e;
        S += k;
    }
    return hipSuccess;
}

}  // namespace bntt
