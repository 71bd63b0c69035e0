// ntt.hip -- launch-per-call forward NTT: `count` transforms of one size with
// one twiddle table, in one launch (N <= 4096) or two or three launches
// through device memory (N > 4096), or one launch over several workgroups
// joined by a grid barrier (a lone transform of 2^11 .. 2^13 points given a
// tmp).  The network and its blocking are described in
// ntt_core.hpp.
#include "internal.hpp"
#include "ntt_core.hpp"

#include <chrono>
#include <random>

namespace bntt {

// One pass: stages [S, S+k) of `count` transforms of size 2^L.
// Group id gamma = transform * 2^(L-k) + M; this workgroup owns groups
// [c * G, c * G + G) for its chunk c = xcd_chunk(blockIdx.x) (each XCD takes
// a contiguous range: 8192 x 1024 1.5% faster, one 2^20 / 2^24 transform 7% /
// 3.5%, measured alternately on MI355X).  Block size = (G << k) >> ELOG.
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
    const size_t gamma0 = xcd_chunk(blockIdx.x, gridDim.x) << Glog;
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
        // a launch whose completion the host reads from pinned memory: every
        // wave makes its stores visible system-wide, then one release store
        // posts the sequence number to this workgroup's word
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(done + blockIdx.x, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Lone transforms of 2^L points, L <= 13: one thread per butterfly of a
// stage, so a stage costs one butterfly's dependency chain rather than a CU's
// share of the whole network.  A transform of up to 128 points is one group of
// one pass; larger ones take two passes of K1 and L - K1 stages over groups of
// at most 64 points.  A group's threads sit in one wavefront, so its stages
// exchange values through the wavefront's own LDS slots with no workgroup
// barrier (LDS executes a wavefront's accesses in order; `lds_order` keeps the
// compiler from moving them).  Between the passes every value crosses once,
// stored in the second pass's group order: through LDS when one workgroup
// holds the transform (P = 1, up to 2^10 points, one workgroup per transform
// of the launch), or through device memory and a grid barrier when a lone
// transform of 2^11, 2^12 or 2^13 points is split over P = 4, 8 or 16
// workgroups of 256 threads.  (At 2^10 points two workgroups measured slower
// than one: 8.1 against 7.3 us per call, the barrier and the round trip
// through device memory costing more than the halved network saves.  Half as
// many workgroups of 512 threads were slower at 2^11 .. 2^14: 10.8 against
// 8.6 us at 2^11, 11.8 against 10.1 us at 2^13; and 2^14 over 32 workgroups,
// 13.2 us, did not beat its two launches, 13.0 us.)  Every input and the
// one twiddle of each of the thread's L stages are loaded before the first
// butterfly, and all loads precede every store (y may equal x).
//
// Thread mapping: in pass 1 the groups of a wavefront interleave (group g of
// the wavefront is lane g + GPW1 * i), so consecutive lanes load consecutive
// residues and write the intermediate contiguously; in pass 2 each group's
// lanes are consecutive, so they read the intermediate and store y
// contiguously.  The arithmetic is that of the other kernels: the fast exact
// product on a wavefront whose inputs are at most q and whose twiddles are
// below 2^124 (decided per wavefront and pass), the literal reference
// formulas otherwise.
__device__ __forceinline__ void lds_order() { __asm__ volatile("" ::: "memory"); }

template <int L, int P>
struct LoneShape {
    static constexpr int T = (1 << (L - 1)) / P;                // threads = butterflies per stage per workgroup
    static constexpr int W = T < 64 ? T : 64;                   // threads per wavefront
    static constexpr int NW = T / W;                            // wavefronts
    static constexpr int K1 = (P == 1 && L <= 7) ? L : L / 2;   // stages of pass 1
    static constexpr int K2 = L - K1;                           // stages of pass 2 (0: one pass)
    static constexpr int G1 = 1 << (L - K1);                    // groups of pass 1
    static constexpr int LPG1 = 1 << (K1 - 1);                  // threads per group, pass 1
    static constexpr int GPW1 = W / LPG1;                       // groups per wavefront, pass 1
    static constexpr int LPG2 = K2 ? 1 << (K2 - 1) : 1;         // threads per group, pass 2
    static constexpr int GPW2 = W / LPG2;                       // groups per wavefront, pass 2
    static_assert(LPG1 <= W && LPG2 <= W, "a group's threads must share a wavefront");
};

// Grid-barrier polls before a workgroup gives up on its siblings: each poll is
// a device-scope load of about 0.5-1 us, so about one to two seconds (every
// workgroup of this library's kernels is resident far sooner).
constexpr uint32_t kLoneSpinLimit = 1u << 21;

// One stage of a group: p = tw * b, (a, b) <- (a + p, a - p) as the
// reference's ntt_scalar.c:611-618.  FAST: digit twiddle, fast exact product.
template <bool FAST, bool PM>
__device__ __forceinline__ void lone_butterfly(U128 &a, U128 &b, const R31 &t, const U128 *tw_entry, const Mod &md)
{
    if constexpr (FAST) {
        const U128 p = mulmod_r31(b, t, md.c);
        b = submod_pm(a, p, md.m.v[0]);
        a = addmod_pm(a, p, md.m.v[0]);
    } else {
        const U128 p = mulmod<PM>(ld16(tw_entry), b, md);
        b = submod(a, p, md);
        a = addmod(a, p, md);
    }
}

// K stages of the wavefront's groups.  Thread (group g, butterfly j) holds
// the group's elements j and j + 2^(K-1); after a stage it holds outputs 2j
// and 2j+1, which go through the wavefront's slots to the threads that need
// them next: value v of the thread's group sits at xw[v * slot_stride +
// slot_off].  tw_idx[r]: this thread's entry of the caller's table at stage r
// (read again only on the literal path).
template <int K, bool FAST, bool PM, bool SWZ = false>
__device__ __forceinline__ void lone_group_stages(U128 &a, U128 &b, const R31 *t, const uint32_t *tw_idx,
                                                  const U128 *__restrict__ tw, U128 *xw, int slot_stride,
                                                  int slot_off, int j, const Mod &md, int half = -1)
{
    constexpr int H = 1 << (K - 1);
    // SWZ (pass 2, whose lanes hold consecutive butterflies of one group):
    // the writes of outputs 2j and 2j + 1 step by two slots, which puts two
    // lanes of every 8 (or 16) on one LDS bank group; flipping bit 0 of the
    // slot by bits 3 and 4 spreads them, and the reads of j and j + H stay
    // conflict-free (a bijection on the wavefront's slots, the same for the
    // writes and reads of a stage)
    auto slot = [&](int v) {
        const int s = v * slot_stride + slot_off;
        return SWZ ? s ^ (((s >> 3) ^ (s >> 4)) & 1) : s;
    };
#pragma unroll
    for (int r = 0; r < K; r++) {
        // half >= 0: the two wavefronts that share a SIMD (half 0 and 1; the
        // workgroup's eight wavefronts go to the CU's four SIMDs in turn, so
        // wavefronts w and w + 4 pair up) take turns at the higher issue
        // priority, one stage each, the older one first.  They then finish the
        // pass together, rather than the older one first and the younger one
        // alone, at one wavefront's slower pace (measured: 6.81 -> 6.70 us per
        // call at 2^10, 7.29 -> 6.95 us for 64 transforms; the older one first
        // rather than the younger, 6.63 -> 6.55 and 6.90 -> 6.82 us).
        if (half >= 0) {
            if (((r + half + 1) & 1) != 0)
                __builtin_amdgcn_s_setprio(1);
            else
                __builtin_amdgcn_s_setprio(0);
        }
        lone_butterfly<FAST, PM>(a, b, t[r], tw + tw_idx[r], md);
        if (r + 1 < K) {
            xw[slot(2 * j)] = a;
            xw[slot(2 * j + 1)] = b;
            lds_order();
            a = xw[slot(j)];
            b = xw[slot(j + H)];
            lds_order();
        }
    }
}

// zg, bar, token (P > 1): the intermediate (2^L residues) and P barrier words
// of this launch, and the launch's token, which no earlier launch on these
// words used.  done (host-array calls): workgroup w posts seq to done[w] once
// its outputs are visible system-wide, or seq ^ 0x80000000 if its barrier
// gave up (its outputs are then wrong).
template <int L, int P, bool PM>
__global__ __launch_bounds__((LoneShape<L, P>::T)) void ntt_lone_kernel(
    U128 *__restrict__ y, const U128 *__restrict__ x, const U128 *__restrict__ tw, Mod md, U128 *__restrict__ zg,
    uint64_t *bar, uint64_t token, uint32_t *done, uint32_t seq)
{
    using S = LoneShape<L, P>;
    constexpr int N = 1 << L;
    constexpr int K1 = S::K1, K2 = S::K2;
    // per wavefront 2W exchange slots; for two passes in one workgroup also
    // the intermediate
    __shared__ U128 xch[S::NW * 2 * S::W];
    __shared__ U128 z[(K2 && P == 1) ? N : 1];
    const int t = threadIdx.x;
    const int wv = t / S::W, l = t % S::W;
    U128 *xw = xch + wv * 2 * S::W;
    // P = 1: workgroup b does transform b; P > 1: part b of the one transform
    const int part = P == 1 ? 0 : (int)blockIdx.x;
    if constexpr (P == 1) {
        x += (size_t)blockIdx.x * N;
        y += (size_t)blockIdx.x * N;
    }

    // pass 1: group M1, butterfly j1; pass 2: group M2, butterfly j2
    const int g1 = l % S::GPW1, j1 = l / S::GPW1;
    const int M1 = (part * S::NW + wv) * S::GPW1 + g1;
    const int j2 = l % S::LPG2, g2 = l / S::LPG2;
    const int M2 = (part * S::NW + wv) * S::GPW2 + g2;

    U128 a = ld16(x + j1 * S::G1 + M1), b = ld16(x + (j1 + S::LPG1) * S::G1 + M1);
    uint32_t idx[L];
    U128 traw[L];
#pragma unroll
    for (int r = 0; r < K1; r++) idx[r] = (1u << r) + ((uint32_t)j1 & ((1u << r) - 1u));
#pragma unroll
    for (int r = 0; r < K2; r++) idx[K1 + r] = (1u << (K1 + r)) + ((uint32_t)M2 << r) + ((uint32_t)j2 & ((1u << r) - 1u));
#pragma unroll
    for (int r = 0; r < L; r++) traw[r] = ld16(tw + idx[r]);
    bool big1 = false, big2 = false;
    R31 tr[L];
#pragma unroll
    for (int r = 0; r < L; r++) {
        (r < K1 ? big1 : big2) |= tw_big(traw[r]);
        tr[r] = to_r31(traw[r]);
    }

    // two wavefronts per SIMD (eight in the workgroup): alternate their priority
    const int half = S::NW == 8 ? (wv >> 2) : -1;
    const bool slow1 = !PM || __any(big1 || above_q(a, md.m) || above_q(b, md.m));
    if (!slow1) {
        if constexpr (PM) lone_group_stages<K1, true, PM>(a, b, tr, idx, tw, xw, S::GPW1, g1, j1, md, half);
    } else {
        lone_group_stages<K1, false, PM>(a, b, tr, idx, tw, xw, S::GPW1, g1, j1, md);
    }
    bool timed_out = false;
    if constexpr (K2 == 0) {
        // one group: outputs 2j and 2j+1 are consecutive
        st16(y + 2 * j1, a);
        st16(y + 2 * j1 + 1, b);
    } else {
        // output i of pass-1 group M at Z[i * G1 + M]: pass-2 group M2 reads
        // its inputs Z[M2 * G1 + i] contiguously
        if constexpr (P == 1) {
            z[(2 * j1) * S::G1 + M1] = a;
            z[(2 * j1 + 1) * S::G1 + M1] = b;
            __syncthreads();
            a = z[M2 * S::G1 + j2];
            b = z[M2 * S::G1 + j2 + S::LPG2];
        } else {
            st16(zg + (2 * j1) * S::G1 + M1, a);
            st16(zg + (2 * j1 + 1) * S::G1 + M1, b);
            // grid barrier: every wavefront releases its stores at device
            // scope, then one thread posts this workgroup's word and the first
            // P threads wait for all P words to carry this launch's token
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (t == 0) __hip_atomic_store(&bar[part], token, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (t < P) {
                uint32_t n = 0;
                while (__hip_atomic_load(&bar[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != token)
                    if (++n == kLoneSpinLimit) {
                        timed_out = true;
                        break;
                    }
            }
            timed_out = __syncthreads_or(timed_out);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            a = ld16(zg + M2 * S::G1 + j2);
            b = ld16(zg + M2 * S::G1 + j2 + S::LPG2);
        }
        const bool slow2 = !PM || __any(big2 || above_q(a, md.m) || above_q(b, md.m));
        if (!slow2) {
            if constexpr (PM)
                lone_group_stages<K2, true, PM, true>(a, b, tr + K1, idx + K1, tw, xw, 1, g2 << K2, j2, md, half);
        } else {
            lone_group_stages<K2, false, PM, true>(a, b, tr + K1, idx + K1, tw, xw, 1, g2 << K2, j2, md);
        }
        st16(y + (M2 << K2) + 2 * j2, a);
        st16(y + (M2 << K2) + 2 * j2 + 1, b);
    }
    if (done) {
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(done + part, timed_out ? seq ^ 0x80000000u : seq, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Grid-barrier tokens: one counter for every split launch of the process,
// from a random start, so that no launch's token equals a stale barrier word
// that an earlier launch (of any size, in this process or an earlier one)
// left in the same memory.
static uint64_t next_barrier_token()
{
    static std::atomic<uint64_t> next{((uint64_t)std::random_device{}() << 32) ^
                                      (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count()};
    return next.fetch_add(1) + 1;
}

template <int L, int P, bool PM>
static hipError_t launch_lone_l(U128 *y, const U128 *x, const U128 *tw, size_t count, const Mod &md, U128 *tmp,
                                hipStream_t stream, uint32_t *done, uint32_t seq)
{
    U128 *zg = nullptr;
    uint64_t *bar = nullptr;
    uint64_t token = 0;
    if constexpr (P > 1) {
        zg = tmp;
        bar = reinterpret_cast<uint64_t *>(tmp + ((size_t)1 << L));
        token = next_barrier_token();
        count = P;
    }
    hipLaunchKernelGGL((ntt_lone_kernel<L, P, PM>), dim3((unsigned)count), dim3(LoneShape<L, P>::T), 0, stream, y, x,
                       tw, md, zg, bar, token, done, seq);
    return hipGetLastError();
}

template <bool PM>
static hipError_t launch_lone(U128 *y, const U128 *x, const U128 *tw, int L, size_t count, const Mod &md,
                              U128 *tmp, hipStream_t stream, uint32_t *done, uint32_t seq)
{
    if (tmp) {
        switch (L) {
        case 11: return launch_lone_l<11, 4, PM>(y, x, tw, 1, md, tmp, stream, done, seq);
        case 12: return launch_lone_l<12, 8, PM>(y, x, tw, 1, md, tmp, stream, done, seq);
        case 13: return launch_lone_l<13, 16, PM>(y, x, tw, 1, md, tmp, stream, done, seq);
        default: return hipErrorInvalidValue;
        }
    }
    switch (L) {
    case 1: return launch_lone_l<1, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 2: return launch_lone_l<2, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 3: return launch_lone_l<3, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 4: return launch_lone_l<4, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 5: return launch_lone_l<5, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 6: return launch_lone_l<6, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 7: return launch_lone_l<7, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 8: return launch_lone_l<8, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 9: return launch_lone_l<9, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    case 10: return launch_lone_l<10, 1, PM>(y, x, tw, count, md, nullptr, stream, done, seq);
    default: return hipErrorInvalidValue;
    }
}

template <int ELOG, bool PM>
static hipError_t launch_pass(U128 *y, const U128 *x, const U128 *tw, int L, int S, int k, int Glog, size_t count,
                              const Mod &md, hipStream_t stream, uint32_t *done = nullptr, uint32_t seq = 0,
                              int *nblocks = nullptr, int done_cap = 0)
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
    if (done && blocks > (size_t)done_cap) return hipErrorInvalidValue;  // a word per workgroup
    if (nblocks) *nblocks = (int)blocks;
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
                                uint32_t seq = 0, int *nblocks = nullptr, int done_cap = 0)
{
    switch (elog) {
    case 1:
        return launch_pass<1, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq, nblocks, done_cap);
    case 2:
        return launch_pass<2, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq, nblocks, done_cap);
    default:
        return launch_pass<3, PM>(y, x, tw, L, S, k, Glog, count, md, stream, done, seq, nblocks, done_cap);
    }
}

// Transforms per launch below which the single-pass kernel gives each
// transform its own workgroup of up to 512 threads (latency-bound regime)
// instead of packing 2048 values of several transforms per workgroup.
constexpr size_t kFewTransforms = 64;
// largest lone transform on one workgroup with one thread per butterfly of a
// stage (512 threads)
constexpr int kLoneMaxLog2 = 10;

hipError_t launch_ntt_uniform(U128 *y, const U128 *x, const U128 *tw, int L, size_t count, const Mod &md, U128 *tmp,
                              hipStream_t stream, uint32_t *done, uint32_t seq, int *done_words, int done_cap)
{
    if (done_words) *done_words = 1;
    if (count == 0) return hipSuccess;
    if (L < kMinLog2 || L > kMaxLog2) return hipErrorInvalidValue;
    // completion words: one transform of at most 2^13 points in one launch,
    // each of whose workgroups posts one (*done_words of them, at most
    // done_cap)
    if (done && (count != 1 || L > kLoneSplitMaxLog2 || done_cap < 1)) return hipErrorInvalidValue;
    const bool pm = md.pm != 0;
    // A lone transform of 2^11 .. 2^13 points given a tmp splits over several
    // workgroups joined by a grid barrier.  Its per-launch token would repeat
    // in every replay of a captured graph, so a launch being captured takes
    // the paths below instead.
    if (tmp != nullptr && count == 1 && L >= kLoneSplitMinLog2 && L <= kLoneSplitMaxLog2) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        const hipError_t ce = hipStreamIsCapturing(stream, &cs);
        if (ce != hipSuccess) return ce;
        if (cs == hipStreamCaptureStatusNone) {
            if (done && lone_parts(L) > done_cap) return hipErrorInvalidValue;
            if (done_words) *done_words = lone_parts(L);
            return pm ? launch_lone<true>(y, x, tw, L, 1, md, tmp, stream, done, seq)
                      : launch_lone<false>(y, x, tw, L, 1, md, tmp, stream, done, seq);
        }
    }
    if (L <= kLoneMaxLog2 && count <= kFewTransforms)
        return pm ? launch_lone<true>(y, x, tw, L, count, md, nullptr, stream, done, seq)
                  : launch_lone<false>(y, x, tw, L, count, md, nullptr, stream, done, seq);
    // A lone transform of 2^11 or 2^12 points on one workgroup is
    // latency-bound (15 and 26 us); given a tmp under capture, it takes two
    // launches over several CUs instead.
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
        return pm ? launch_pass_e<true>(y, x, tw, L, 0, L, Glog, elog, count, md, stream, done, seq, done_words,
                                        done_cap)
                  : launch_pass_e<false>(y, x, tw, L, 0, L, Glog, elog, count, md, stream, done, seq, done_words,
                                         done_cap);
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
