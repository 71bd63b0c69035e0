// internal.hpp -- host helpers shared by the C ABI translation units
// (hidden symbols; the exported entry points are those of include/bntt.h).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "../../include/bntt.h"
#include "u128.hpp"

// Record a failure for bntt_last_error() and return `code`.
int bntt_set_error(int code, const char *msg);
int bntt_set_hip_error(hipError_t e, const char *what);
// Check the current device is gfx950 (once per device).
int bntt_ensure_init();
// Modulus descriptor, including the check that selects the fast reduction.
bntt::Mod bntt_make_mod(const bntt_u128 *m, const bntt_u128 *mu);

// Raise a kernel's dynamic LDS limit to `bytes` on the current device, once per
// device (`done` keeps one bit per device index below 64, set on success).
inline hipError_t bntt_allow_lds(const void *kernel, int bytes, std::atomic<uint64_t> &done)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = dev < 64 ? (1ull << dev) : 0;
    if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess) done.fetch_or(bit, std::memory_order_release);
    return e;
}
