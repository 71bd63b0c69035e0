// ntt.hip -- forward constant-geometry (Pease) 128-bit NTT for gfx950.
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
// The same split applies recursively inside a group, so one kernel does:
//   HBM -> LDS (G groups, coalesced) -> register sub-passes of up to 2^ELOG
//   points per thread (LDS -> VGPR -> LDS) -> LDS -> HBM (contiguous).
// Transforms of up to 4096 points take one such pass (all L stages, several
// transforms per workgroup for small N); larger ones take ceil(L/10) passes
// through HBM.  All twiddle indices are computed, never stored.
#include "kernels.hpp"

namespace bntt {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ U128 ld16(const U128 *p)
{
    const v4u t = *reinterpret_cast<const v4u *>(p);
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
    *reinterpret_cast<v4u *>(p) = t;
}

constexpr int kMaxThreads = 512;

// KP stages of the constant-geometry network on 2^KP values held in registers,
// starting at global stage s0 with twiddle prefix P (s0 bits).
template <int KP, bool PM>
__device__ __forceinline__ void reg_network(U128 (&v)[1 << KP], const U128 *__restrict__ tw, int s0, uint32_t P,
                                            const Mod &md)
{
    constexpr int H = 1 << (KP - 1);
#pragma unroll
    for (int r = 0; r < KP; r++) {
        U128 t[H];
        const uint32_t base = (1u << (s0 + r)) + (P << r);
#pragma unroll
        for (int j = 0; j < (1 << r) && j < H; j++) t[j] = ld16(tw + base + j);
        U128 w[1 << KP];
#pragma unroll
        for (int i = 0; i < H; i++) {
            const U128 p = mulmod<PM>(t[i & ((1 << r) - 1)], v[i + H], md);
            w[2 * i] = addmod(v[i], p, md);
            w[2 * i + 1] = submod(v[i], p, md);
        }
#pragma unroll
        for (int i = 0; i < 2 * H; i++) v[i] = w[i];
    }
}

// One register sub-pass of KP stages over the G groups of 2^k values in LDS.
// Each thread owns 2^(ELOG-KP) sub-groups.  r0: local stage offset inside the
// group network; the group's global stage offset is S and its prefix P0[g].
template <int ELOG, int KP, bool PM>
__device__ __forceinline__ void lds_subpass(U128 *lds, const U128 *__restrict__ tw, int k, int S, int r0,
                                            uint32_t mbase, const Mod &md)
{
    constexpr int Q = 1 << (ELOG - KP);
    constexpr int R = 1 << KP;
    const int T = blockDim.x;
    const int sub_log = k - KP;  // sub-groups per group = 2^sub_log
    U128 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = threadIdx.x + q * T;
        const int g = sg >> sub_log;
        const int Mp = sg & ((1 << sub_log) - 1);
        const U128 *src = lds + (g << k) + Mp;
#pragma unroll
        for (int j = 0; j < R; j++) v[q][j] = src[j << sub_log];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int sg = threadIdx.x + q * T;
        const int g = sg >> sub_log;
        const int Mp = sg & ((1 << sub_log) - 1);
        // group prefix: low S bits of the group's M, then low r0 bits of Mp
        const uint32_t P0 = (mbase + (uint32_t)g) & ((1u << S) - 1u);
        const uint32_t P = (P0 << r0) | ((uint32_t)Mp & ((1u << r0) - 1u));
        reg_network<KP, PM>(v[q], tw, S + r0, P, md);
        U128 *dst = lds + (g << k) + (Mp << KP);
#pragma unroll
        for (int j = 0; j < R; j++) dst[j] = v[q][j];
    }
    __syncthreads();
}

// One pass: stages [S, S+k) of `count` transforms of size 2^L.
// Group id gamma = transform * 2^(L-k) + M; this workgroup owns groups
// [blockIdx.x * G, blockIdx.x * G + G).  Block size = (G << k) >> ELOG.
template <int ELOG, bool PM>
__global__ __launch_bounds__(kMaxThreads) void ntt_pass_kernel(U128 *__restrict__ y, const U128 *__restrict__ x,
                                                               const U128 *__restrict__ tw, int L, int S, int k,
                                                               int Glog, size_t total_groups, Mod md)
{
    extern __shared__ U128 lds[];
    constexpr int E = 1 << ELOG;
    const int T = blockDim.x;
    const int G = 1 << Glog;
    const int mlog = L - k;  // groups per transform = 2^mlog
    const size_t gamma0 = (size_t)blockIdx.x << Glog;
    const int nvalid = (int)((total_groups - gamma0) < (size_t)G ? (total_groups - gamma0) : (size_t)G);

    // HBM -> LDS.  Single-pass (k == L): groups are whole transforms and the
    // block's data is one contiguous range.  Otherwise element T_idx of group g
    // sits at transform*N + T_idx*2^mlog + M; consecutive threads take
    // consecutive g (consecutive addresses).
#pragma unroll
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * T;
        int g, ti;
        if (mlog == 0) {
            g = e >> k;
            ti = e & ((1 << k) - 1);
        } else {
            g = e & (G - 1);
            ti = e >> Glog;
        }
        if (g < nvalid) {
            const size_t gamma = gamma0 + (size_t)g;
            const size_t tr = gamma >> mlog;
            const size_t M = gamma & (((size_t)1 << mlog) - 1);
            lds[(g << k) + ti] = ld16(x + (tr << L) + ((size_t)ti << mlog) + M);
        }
    }
    __syncthreads();

    const uint32_t mbase = (uint32_t)(gamma0 & (((size_t)1 << mlog) - 1));
    int r0 = 0;
    while (r0 < k) {
        const int kp = (k - r0) < ELOG ? (k - r0) : ELOG;
        if (kp == ELOG) {
            lds_subpass<ELOG, ELOG, PM>(lds, tw, k, S, r0, mbase, md);
        } else if constexpr (ELOG >= 2) {
            if (kp == 1) {
                lds_subpass<ELOG, 1, PM>(lds, tw, k, S, r0, mbase, md);
            } else if constexpr (ELOG >= 3) {
                if (kp == 2) lds_subpass<ELOG, 2, PM>(lds, tw, k, S, r0, mbase, md);
            }
        }
        r0 += kp;
    }

    // LDS -> HBM: the G groups' outputs are one contiguous range.
    const size_t out0 = gamma0 << k;
#pragma unroll
    for (int j = 0; j < E; j++) {
        const int e = threadIdx.x + j * T;
        if ((e >> k) < nvalid) st16(y + out0 + e, lds[e]);
    }
}

template <int ELOG, bool PM>
static hipError_t launch_pass(U128 *y, const U128 *x, const U128 *tw, int L, int S, int k, int Glog, size_t count,
                              const Mod &md, hipStream_t stream)
{
    const size_t total_groups = count << (L - k);
    const size_t blocks = (total_groups + ((size_t)1 << Glog) - 1) >> Glog;
    const int threads = (1 << (Glog + k)) >> ELOG;
    const size_t lds_bytes = sizeof(U128) << (Glog + k);
    if (threads < 1 || threads > kMaxThreads || blocks > 0x7fffffffu || lds_bytes > 65536)
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((ntt_pass_kernel<ELOG, PM>), dim3((unsigned)blocks), dim3(threads), lds_bytes, stream, y, x,
                       tw, L, S, k, Glog, total_groups, md);
    return hipGetLastError();
}

template <bool PM>
static hipError_t launch_pass_e(U128 *y, const U128 *x, const U128 *tw, int L, int S, int k, int Glog,
                                size_t count, const Mod &md, hipStream_t stream)
{
    const int elog = k < 3 ? k : 3;
    switch (elog) {
    case 1:
        return launch_pass<1, PM>(y, x, tw, L, S, k, Glog, count, md, stream);
    case 2:
        return launch_pass<2, PM>(y, x, tw, L, S, k, Glog, count, md, stream);
    default:
        return launch_pass<3, PM>(y, x, tw, L, S, k, Glog, count, md, stream);
    }
}

hipError_t launch_ntt_uniform(U128 *y, const U128 *x, const U128 *tw, int L, size_t count, const Mod &md, U128 *tmp,
                              hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    if (L < kMinLog2 || L > kMaxLog2) return hipErrorInvalidValue;
    const bool pm = md.pm != 0;
    if (L <= kLdsMaxLog2) {
        // one pass, several transforms per workgroup: 256 threads of up to 8 points
        const int elog = L < 3 ? L : 3;
        int Glog = 8 + elog - L;
        if (Glog < 0) Glog = 0;
        return pm ? launch_pass_e<true>(y, x, tw, L, 0, L, Glog, count, md, stream)
                  : launch_pass_e<false>(y, x, tw, L, 0, L, Glog, count, md, stream);
    }
    if (tmp == nullptr) return hipErrorInvalidValue;
    const int np = (L + 9) / 10;
    int S = 0;
    for (int p = 0; p < np; p++) {
        const int k = (L - S) / (np - p);
        const int Glog = 12 - k;  // 4096 values per workgroup
        const U128 *src = (p == 0) ? x : ((np - p) % 2 == 0 ? y : tmp);
        U128 *dst = ((np - 1 - p) % 2 == 0) ? y : tmp;
        const hipError_t e = pm ? launch_pass_e<true>(dst, src, tw, L, S, k, Glog, count, md, stream)
                                : launch_pass_e<false>(dst, src, tw, L, S, k, Glog, count, md, stream);
        if (e != hipSuccess) return e;
        S += k;
    }
    return hipSuccess;
}

}  // namespace bntt
