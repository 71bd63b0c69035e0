// io.hip -- the input side of the path: conversion between the hi/lo word
// order of the reference's fixture files and AVX code (config.c:6-197,
// ntt_avx512.c:386-528) and native u128; twiddle tables of roots of unity;
// seeded uniform residues; and the reference's scalar primitives
// (ntt_scalar.c:49-199) as host functions.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bntt.h"
#include "hostmont.hpp"
#include "internal.hpp"

using bntt::HostMont;
using bntt::MontQ;
using bntt::R31;
using bntt::U128;

namespace {

typedef unsigned __int128 u128;

constexpr int kIoThreads = 256;

int grid_for(size_t n)
{
    const size_t b = (n + kIoThreads - 1) / kIoThreads;
    return (int)(b < 8192 ? (b ? b : 1) : 8192);
}

// (w0, w1) -> (w1, w0) per 16-byte element; out may equal in.
__global__ __launch_bounds__(kIoThreads) void swap_words_kernel(ulonglong2 *out, const ulonglong2 *in, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kIoThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kIoThreads) {
        const ulonglong2 v = in[i];
        out[i] = make_ulonglong2(v.y, v.x);
    }
}

// Root-of-unity table: entry 2^s + k (s < L, k < 2^s) = root[s]^e(k) with
// e(k) = brv_s(k) (cyclic) or 2 brv_s(k) + 1 (negacyclic); entry 0 = 0.
struct TwGen {
    MontQ mq;
    U128 one;        // Montgomery form of 1
    U128 root[24];   // Montgomery form of the base root of stage s
    int log2n;
    int nega;
};

__global__ __launch_bounds__(kIoThreads) void twiddle_kernel(U128 *tw, TwGen g)
{
    const size_t n = (size_t)1 << g.log2n;
    for (size_t e = blockIdx.x * (size_t)kIoThreads + threadIdx.x; e < n; e += (size_t)gridDim.x * kIoThreads) {
        if (e == 0) {
            tw[0] = bntt::make_u128(0, 0);
            continue;
        }
        const int s = 31 - __clz((uint32_t)e);
        const uint32_t k = (uint32_t)e - (1u << s);
        uint32_t ex = s ? (__brev(k) >> (32 - s)) : 0u;
        if (g.nega) ex = 2 * ex + 1;
        U128 acc = g.one, b = g.root[s];
        while (ex) {
            if (ex & 1) acc = bntt::montmul_r31(acc, bntt::to_r31(b), g.mq);
            ex >>= 1;
            if (ex) b = bntt::montmul_r31(b, bntt::to_r31(b), g.mq);
        }
        R31 unit{};
        unit.d[0] = 1;
        tw[e] = bntt::montmul_r31(acc, unit, g.mq);  // leave the Montgomery domain
    }
}

// Word j of the splitmix64 sequence started at `seed` (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014; the
// finaliser constants of Vigna's public-domain splitmix64.c).
__host__ __device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t j)
{
    uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Element i, attempt t draws words 2(t n + i) (low word) and 2(t n + i) + 1
// (high word), masked below 2^bitlen(q); the first draw below q is kept.  The
// 64th attempt, if reached, keeps its draw minus q (< q as the draw < 2q).
__global__ __launch_bounds__(kIoThreads) void uniform_kernel(ulonglong2 *out, size_t n, uint64_t ql, uint64_t qh,
                                                             uint64_t lomask, uint64_t himask, uint64_t seed)
{
    for (size_t i = blockIdx.x * (size_t)kIoThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kIoThreads) {
        uint64_t lo = 0, hi = 0;
        for (int t = 0; t < 64; t++) {
            const uint64_t j = 2 * ((uint64_t)t * n + i);
            lo = splitmix_at(seed, j) & lomask;
            hi = splitmix_at(seed, j + 1) & himask;
            if (hi < qh || (hi == qh && lo < ql)) break;
            if (t == 63) {
                hi = hi - qh - (lo < ql);
                lo = lo - ql;
            }
        }
        out[i] = make_ulonglong2(lo, hi);
    }
}

int check_dev_ptr(const void *p, const char *what)
{
    if (!p || (reinterpret_cast<uintptr_t>(p) & 15u)) {
        char buf[96];
        snprintf(buf, sizeof buf, "%s must be non-null and 16-byte aligned", what);
        return bntt_set_error(BNTT_EINVAL, buf);
    }
    return BNTT_OK;
}

int launched(const char *what)
{
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? BNTT_OK : bntt_set_hip_error(e, what);
}

int make_twgen(TwGen *g, int log2n, const bntt_u128 *q, int kind)
{
    if (!q) return bntt_set_error(BNTT_EINVAL, "null modulus");
    if (log2n < 1 || log2n > 24) return bntt_set_error(BNTT_EINVAL, "log2n must be in 1..24");
    if (kind < BNTT_TW_CYCLIC || kind > BNTT_TW_NEGACYCLIC_INV) return bntt_set_error(BNTT_EINVAL, "bad table kind");
    HostMont hm;
    if (!hm.init(((u128)q->hi << 64) | q->lo)) return bntt_set_error(BNTT_EINVAL, "q must be odd, >= 3 and < 2^122");
    if (!hm.is_probable_prime()) return bntt_set_error(BNTT_EINVAL, "q is not prime");
    U128 w{};
    int v = 0;
    if (!hm.two_adic_root(&w, &v)) return bntt_set_error(BNTT_EINVAL, "no quadratic non-residue below 1000");
    const bool nega = kind == BNTT_TW_NEGACYCLIC || kind == BNTT_TW_NEGACYCLIC_INV;
    const bool inv = kind == BNTT_TW_CYCLIC_INV || kind == BNTT_TW_NEGACYCLIC_INV;
    if (log2n + (nega ? 1 : 0) > v) return bntt_set_error(BNTT_EINVAL, "q - 1 has too few factors of two for this size");
    g->mq = hm.mq;
    g->one = hm.r1;
    g->log2n = log2n;
    g->nega = nega;
    for (int s = 0; s < log2n; s++) {
        const int order = s + (nega ? 2 : 1);  // log2 of the stage root's order
        U128 r = hm.pow(w, (u128)1 << (v - order));
        if (inv) r = hm.pow(r, ((u128)1 << order) - 1);
        g->root[s] = r;
    }
    return BNTT_OK;
}

int uniform_args(const bntt_u128 *q, uint64_t *lomask, uint64_t *himask)
{
    if (!q || (q->hi == 0 && q->lo < 2)) return bntt_set_error(BNTT_EINVAL, "q must be at least 2");
    const int bits = q->hi ? 128 - __builtin_clzll(q->hi) : 64 - __builtin_clzll(q->lo);
    *lomask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    *himask = bits <= 64 ? 0ull : (bits == 128 ? ~0ull : ((1ull << (bits - 64)) - 1));
    return BNTT_OK;
}

// Run `fill` into a temporary device buffer of n residues, copy to host.
template <class F>
int via_device(bntt_u128 *host, size_t n, F fill)
{
    if (!host && n) return bntt_set_error(BNTT_EINVAL, "null output");
    if (!n) return BNTT_OK;
    if (n > SIZE_MAX / sizeof(bntt_u128)) return bntt_set_error(BNTT_EINVAL, "n is too large");
    void *d = nullptr;
    if (hipMalloc(&d, n * sizeof(bntt_u128)) != hipSuccess) return bntt_set_error(BNTT_ENOMEM, "hipMalloc failed");
    int rc = fill(d);
    if (!rc) {
        const hipError_t e = hipMemcpy(host, d, n * sizeof(bntt_u128), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = bntt_set_hip_error(e, "hipMemcpy");
    }
    (void)hipFree(d);
    return rc;
}

inline U128 lim(const bntt_u128 &v) { return bntt::make_u128(v.lo, v.hi); }
inline bntt_u128 unlim(const U128 &v)
{
    bntt_u128 r;
    r.lo = bntt::lo64(v);
    r.hi = bntt::hi64(v);
    return r;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

bntt_u128 bntt_addmod128(bntt_u128 a, bntt_u128 b, bntt_u128 m)
{
    const U128 mm = lim(m);
    U128 notm;
    for (int i = 0; i < 4; i++) notm.v[i] = ~mm.v[i];
    return unlim(bntt::addmod_ref(lim(a), lim(b), mm, notm));
}

bntt_u128 bntt_submod128(bntt_u128 a, bntt_u128 b, bntt_u128 m)
{
    return unlim(bntt::submod_ref(lim(a), lim(b), lim(m)));
}

bntt_u128 bntt_mulmod128(bntt_u128 a, bntt_u128 b, bntt_u128 m, bntt_u128 mu)
{
    return unlim(bntt::mulmod_ref(lim(a), lim(b), lim(m), lim(mu)));
}

int bntt_hilo_to_u128(bntt_u128 *out, const uint64_t *hilo, size_t n)
{
    if (n && (!out || !hilo)) return bntt_set_error(BNTT_EINVAL, "null argument");
    for (size_t i = 0; i < n; i++) {
        const uint64_t hi = hilo[2 * i], lo = hilo[2 * i + 1];
        out[i].lo = lo;
        out[i].hi = hi;
    }
    return BNTT_OK;
}

int bntt_u128_to_hilo(uint64_t *hilo, const bntt_u128 *in, size_t n)
{
    if (n && (!hilo || !in)) return bntt_set_error(BNTT_EINVAL, "null argument");
    for (size_t i = 0; i < n; i++) {
        const uint64_t lo = in[i].lo, hi = in[i].hi;
        hilo[2 * i] = hi;
        hilo[2 * i + 1] = lo;
    }
    return BNTT_OK;
}

int bntt_hilo_to_u128_dev(void *out, const void *hilo, size_t n, void *stream)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (!n) return BNTT_OK;
    if ((rc = check_dev_ptr(out, "out")) || (rc = check_dev_ptr(hilo, "hilo"))) return rc;
    // in place is fine (each element is read and written by one thread); a
    // shifted overlap would race
    const uintptr_t o = reinterpret_cast<uintptr_t>(out), h = reinterpret_cast<uintptr_t>(hilo);
    if (n > SIZE_MAX / sizeof(bntt_u128)) return bntt_set_error(BNTT_EINVAL, "n is too large");
    const size_t bytes = n * sizeof(bntt_u128);
    if (o != h && o < h + bytes && h < o + bytes)
        return bntt_set_error(BNTT_EINVAL, "out must equal the input or not overlap it");
    hipLaunchKernelGGL(swap_words_kernel, dim3(grid_for(n)), dim3(kIoThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<ulonglong2 *>(out), static_cast<const ulonglong2 *>(hilo), n);
    return launched("word-order conversion launch");
}

int bntt_u128_to_hilo_dev(void *hilo, const void *in, size_t n, void *stream)
{
    return bntt_hilo_to_u128_dev(hilo, in, n, stream);  // the same word swap
}

int bntt_twiddles_dev(void *tw, int log2n, const bntt_u128 *q, int kind, void *stream)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if ((rc = check_dev_ptr(tw, "tw"))) return rc;
    TwGen g{};
    if ((rc = make_twgen(&g, log2n, q, kind))) return rc;
    const size_t n = (size_t)1 << log2n;
    hipLaunchKernelGGL(twiddle_kernel, dim3(grid_for(n)), dim3(kIoThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<U128 *>(tw), g);
    return launched("twiddle table launch");
}

int bntt_twiddles(bntt_u128 *tw, int log2n, const bntt_u128 *q, int kind)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (log2n < 1 || log2n > 24) return bntt_set_error(BNTT_EINVAL, "log2n must be in 1..24");
    return via_device(tw, (size_t)1 << log2n, [&](void *d) {
        int r = bntt_twiddles_dev(d, log2n, q, kind, nullptr);
        if (!r && hipDeviceSynchronize() != hipSuccess) r = bntt_set_error(BNTT_EHIP, "twiddle table failed");
        return r;
    });
}

int bntt_uniform_dev(void *out, size_t n, const bntt_u128 *q, uint64_t seed, void *stream)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (!n) return BNTT_OK;
    if ((rc = check_dev_ptr(out, "out"))) return rc;
    uint64_t lomask, himask;
    if ((rc = uniform_args(q, &lomask, &himask))) return rc;
    hipLaunchKernelGGL(uniform_kernel, dim3(grid_for(n)), dim3(kIoThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<ulonglong2 *>(out), n, q->lo, q->hi, lomask, himask, seed);
    return launched("uniform residues launch");
}

int bntt_uniform(bntt_u128 *out, size_t n, const bntt_u128 *q, uint64_t seed)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    return via_device(out, n, [&](void *d) {
        int r = bntt_uniform_dev(d, n, q, seed, nullptr);
        if (!r && hipDeviceSynchronize() != hipSuccess) r = bntt_set_error(BNTT_EHIP, "uniform residues failed");
        return r;
    });
}

}  // extern "C"
#pragma GCC visibility pop
