// blas.hip -- vvadd / vvsub / vvmul / axpy mod q over 128-bit residues.
//
// Reference: blas_scalar.c:578-608 (vector_add, vector_sub, vector_mul, axpy;
// axpy is c = a + x*b, the scalar multiplying b, blas_scalar.c:605).
//
// These are streaming kernels: 48 bytes of HBM traffic per element (two 16-byte
// operands in, one out).  Each lane moves one whole residue per operand with a
// 16-byte vector access (global_load_dwordx4), so a wave covers 1 KiB per
// access.  The shape was chosen by a sweep on MI355X over elements per lane
// (1/2/4/8), block size (128..1024), one-shot versus grid-stride grids and
// temporal versus non-temporal accesses: one element per lane in 512-thread
// blocks with non-temporal loads and ordinary (write-back) stores moved about
// 7.0 TB/s for vvadd, against 6.5 TB/s with non-temporal stores and 5.8 TB/s
// with ordinary loads.  vvmul/axpy add one mulmod per element (about 75 VALU
// ops, 68 VGPRs); with the wave-uniform fast path, axpy moved 6.7-6.9 TB/s in
// 512-thread blocks and 7.0 TB/s in 256-thread blocks (1024: 6.7 TB/s; two
// or four elements per lane: 6.8-6.9 TB/s).  Alternating 64/128/256-thread
// blocks on three boxes, 128 was 0.5-1% ahead of 256 for every op (64: 1%
// behind), so all four use 128.  Chunks are handed out so that each XCD
// streams one contiguous eighth of the arrays (xcd_chunk): on MI355X that took
// axpy / vvadd / vvmul at 2^24 from 122.1 / 120.2 / 124.5 us to 114.1 / 113.5 /
// 114.1 us on one box, and from 115.8 / 116.7 / 115.7 to 113.9 / 113.3 / 115.1
// us on another (same process, alternating; 128 threads still best, 64 equal,
// 256 and 512 slower, two elements per lane 3% slower).
#include "kernels.hpp"

namespace bntt {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

constexpr int kBlasThreads = 128;
constexpr int kPerLane = 1;

__device__ __forceinline__ U128 load_nt(const U128 *p)
{
    const v4u t = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(p));
    U128 r;
    r.v[0] = t.x;
    r.v[1] = t.y;
    r.v[2] = t.z;
    r.v[3] = t.w;
    return r;
}

__device__ __forceinline__ void store(U128 *p, const U128 &a)
{
    v4u t;
    t.x = a.v[0];
    t.y = a.v[1];
    t.z = a.v[2];
    t.w = a.v[3];
    *reinterpret_cast<v4u *>(p) = t;
}

// DONE (host-array calls): workgroup w posts seq to done[w] once its outputs
// are visible system-wide.
template <int OP, bool PM, bool DONE>
__global__ __launch_bounds__(kBlasThreads) void blas_kernel(U128 *__restrict__ c, const U128 *__restrict__ a,
                                                            const U128 *__restrict__ b, U128 x, size_t n, Mod md,
                                                            uint32_t *done, uint32_t seq)
{
    const size_t base = xcd_chunk(blockIdx.x, gridDim.x) * (kBlasThreads * kPerLane) + threadIdx.x;
    U128 av[kPerLane], bv[kPerLane];
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const size_t i = base + (size_t)j * kBlasThreads;
        if (i < n) {
            av[j] = load_nt(a + i);
            bv[j] = load_nt(b + i);
        }
    }
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const size_t i = base + (size_t)j * kBlasThreads;
        if (i < n) {
            U128 r;
            if (OP == OP_VVADD)
                r = addmod(av[j], bv[j], md);
            else if (OP == OP_VVSUB)
                r = submod(av[j], bv[j], md);
            else if (OP == OP_VVMUL)
                r = mulmod_elem<PM>(av[j], bv[j], md);
            else
                r = addmod(av[j], mulmod_elem<PM>(x, bv[j], md), md);
            store(c + i, r);
        }
    }
    if constexpr (DONE) {
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(done + blockIdx.x, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

size_t blas_blocks(size_t n) { return (n + (size_t)kBlasThreads * kPerLane - 1) / ((size_t)kBlasThreads * kPerLane); }

template <int OP, bool PM>
static hipError_t launch_op(U128 *c, const U128 *a, const U128 *b, U128 x, size_t n, const Mod &md,
                            hipStream_t stream, uint32_t *done, uint32_t seq)
{
    const size_t blocks = blas_blocks(n);
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    if (done)
        hipLaunchKernelGGL((blas_kernel<OP, PM, true>), dim3((unsigned)blocks), dim3(kBlasThreads), 0, stream, c, a,
                           b, x, n, md, done, seq);
    else
        hipLaunchKernelGGL((blas_kernel<OP, PM, false>), dim3((unsigned)blocks), dim3(kBlasThreads), 0, stream, c, a,
                           b, x, n, md, nullptr, 0u);
    return hipGetLastError();
}

hipError_t launch_blas(int op, U128 *c, const U128 *a, const U128 *b, U128 x, size_t n, const Mod &md,
                       hipStream_t stream, uint32_t *done, uint32_t seq)
{
    if (n == 0) return hipSuccess;
    const bool pm = md.pm != 0;
    switch (op) {
    case OP_VVADD:
        return launch_op<OP_VVADD, false>(c, a, b, x, n, md, stream, done, seq);
    case OP_VVSUB:
        return launch_op<OP_VVSUB, false>(c, a, b, x, n, md, stream, done, seq);
    case OP_VVMUL:
        return pm ? launch_op<OP_VVMUL, true>(c, a, b, x, n, md, stream, done, seq)
                  : launch_op<OP_VVMUL, false>(c, a, b, x, n, md, stream, done, seq);
    case OP_AXPY:
        return pm ? launch_op<OP_AXPY, true>(c, a, b, x, n, md, stream, done, seq)
                  : launch_op<OP_AXPY, false>(c, a, b, x, n, md, stream, done, seq);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace bntt
