// probe.hip -- issue rate of the integer VALU instructions the NTT and BLAS
// kernels are made of, measured on the current device: the denominator of the
// NTT's roofline.  Each lane runs kChains independent chains of one
// instruction (written as inline assembly so that exactly that instruction
// is issued, kChains * kIters times per lane); the rate is lane-instructions
// per second over a launch of 16 workgroups of 256 threads per CU.
#include <hip/hip_runtime.h>

#include "../../include/bntt.h"
#include "internal.hpp"

namespace {

constexpr int kChains = 8;
constexpr int kIters = 256;
constexpr int kThreads = 256;

template <int OP>
__global__ __launch_bounds__(kThreads) void valu_probe_kernel(uint64_t *out, uint32_t seed)
{
    uint64_t acc[kChains];
    const uint32_t a = seed + threadIdx.x, b = seed * 3u + 1u;
    const uint64_t smask = (uint64_t)seed * 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = a + c;
    for (int it = 0; it < kIters; it++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) {
            if constexpr (OP == BNTT_PROBE_MAD_U64_U32) {
                asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc[c]) : "v"(a), "v"(b) : "vcc");
            } else if constexpr (OP == BNTT_PROBE_ADD_CO_U32) {
                uint32_t x = (uint32_t)acc[c];
                asm volatile("v_add_co_u32 %0, vcc, %1, %0" : "+v"(x) : "v"(a) : "vcc");
                acc[c] = x;
            } else if constexpr (OP == BNTT_PROBE_CNDMASK_B32) {
                uint32_t x = (uint32_t)acc[c];
                asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(x) : "v"(a), "s"(smask));
                acc[c] = x;
            } else if constexpr (OP == BNTT_PROBE_LSHL_ADD_U64) {
                asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(acc[c]) : "v"((uint64_t)a));
            } else {
                uint32_t x = (uint32_t)acc[c];
                asm volatile("v_and_b32 %0, %1, %0" : "+v"(x) : "v"(a));
                acc[c] = x;
            }
        }
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) s += acc[c];
    out[blockIdx.x * (size_t)kThreads + threadIdx.x] = s;
}

template <int OP>
int time_op(int blocks, uint64_t *out, double *rate)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(valu_probe_kernel<OP>, dim3(blocks), dim3(kThreads), 0, nullptr, out, 1u);
        e = hipDeviceSynchronize();
    }
    const int reps = 10;
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(valu_probe_kernel<OP>, dim3(blocks), dim3(kThreads), 0, nullptr, out, (uint32_t)r);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return bntt_set_hip_error(e, "VALU probe");
    *rate = (double)reps * blocks * kThreads * kChains * kIters / (ms * 1e-3);
    return BNTT_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int bntt_probe_valu_rate(int op, double *lane_ops_per_s, double *nominal_lane_ops_per_s)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (!lane_ops_per_s) return bntt_set_error(BNTT_EINVAL, "null output");
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
        return bntt_set_error(BNTT_EHIP, "hipGetDeviceProperties");
    if (nominal_lane_ops_per_s) *nominal_lane_ops_per_s = (double)p.multiProcessorCount * 64.0 * p.clockRate * 1e3;
    const int blocks = p.multiProcessorCount * 16;
    uint64_t *out = nullptr;
    if (hipMalloc(&out, sizeof(uint64_t) * blocks * kThreads) != hipSuccess)
        return bntt_set_error(BNTT_ENOMEM, "hipMalloc failed");
    switch (op) {
    case BNTT_PROBE_MAD_U64_U32: rc = time_op<BNTT_PROBE_MAD_U64_U32>(blocks, out, lane_ops_per_s); break;
    case BNTT_PROBE_ADD_CO_U32: rc = time_op<BNTT_PROBE_ADD_CO_U32>(blocks, out, lane_ops_per_s); break;
    case BNTT_PROBE_CNDMASK_B32: rc = time_op<BNTT_PROBE_CNDMASK_B32>(blocks, out, lane_ops_per_s); break;
    case BNTT_PROBE_LSHL_ADD_U64: rc = time_op<BNTT_PROBE_LSHL_ADD_U64>(blocks, out, lane_ops_per_s); break;
    case BNTT_PROBE_AND_B32: rc = time_op<BNTT_PROBE_AND_B32>(blocks, out, lane_ops_per_s); break;
    default: rc = bntt_set_error(BNTT_EINVAL, "unknown probe");
    }
    (void)hipFree(out);
    return rc;
}

}  // extern "C"
#pragma GCC visibility pop
