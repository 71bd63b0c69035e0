// api.hip -- the C ABI declared in include/bntt.h.
//
// Host-side responsibilities: argument checks, building the modulus
// descriptor (including the check that selects the fast exact reduction),
// error reporting, and the staging buffers of the host-pointer entry points.
// All arithmetic happens in the kernels of blas.hip and ntt.hip; there is no
// CPU path.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/bntt.h"
#include "internal.hpp"
#include "kernels.hpp"

using bntt::Mod;
using bntt::U128;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(BNTT_EHIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

typedef unsigned __int128 u128;

u128 to_u128(const bntt_u128 &a) { return ((u128)a.hi << 64) | a.lo; }

// floor(2^251 / m) for 2^123 < m < 2^124 by restoring division (the quotient
// is below 2^128).
u128 barrett_mu_251(u128 m)
{
    u128 r = 0, q = 0;
    for (int i = 251; i >= 0; i--) {
        // r < m < 2^124, so the shift cannot overflow
        r = (r << 1) | (i == 251 ? 1 : 0);
        q <<= 1;
        if (r >= m) {
            r -= m;
            q |= 1;
        }
    }
    return q;
}

Mod make_mod_uncached(const bntt_u128 *m, const bntt_u128 *mu)
{
    Mod md{};
    md.m = bntt::make_u128(m->lo, m->hi);
    md.mu = bntt::make_u128(mu ? mu->lo : 0, mu ? mu->hi : 0);
    md.notm = bntt::make_u128(~m->lo, ~m->hi);
    md.c = 0;
    md.pm = 0;
    const u128 mm = to_u128(*m);
    const u128 two124 = (u128)1 << 124;
    if (mu && mm > ((u128)1 << 123) && mm < two124 && two124 - mm < 65536) {
        if (barrett_mu_251(mm) == to_u128(*mu)) {
            md.c = (uint32_t)(two124 - mm);
            md.pm = 1;
        }
    }
    return md;
}

// The descriptor of the calling thread's last modulus (the check of mu costs
// a 252-step division, which would otherwise be paid on every call).
Mod make_mod(const bntt_u128 *m, const bntt_u128 *mu)
{
    thread_local bntt_u128 last_m{~0ull, ~0ull}, last_mu{~0ull, ~0ull};
    thread_local bool last_has_mu = false;
    thread_local Mod last{};
    const bool has_mu = mu != nullptr;
    if (m->lo == last_m.lo && m->hi == last_m.hi && has_mu == last_has_mu &&
        (!has_mu || (mu->lo == last_mu.lo && mu->hi == last_mu.hi)))
        return last;
    last = make_mod_uncached(m, mu);
    last_m = *m;
    last_mu = has_mu ? *mu : bntt_u128{0, 0};
    last_has_mu = has_mu;
    return last;
}

// Intermediates of bntt_ntt_fwd_dev_noalloc: a fixed array in device memory,
// one copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: the device address of each GPU's copy comes
// from hipGetSymbolAddress at that device's init.
constexpr size_t kNoallocResidues = BNTT_NOALLOC_MAX_RESIDUES;
static __device__ __align__(256) uint32_t scratch[4 * kNoallocResidues];

// Per device: whether it was checked to be gfx950, and its scratch address.
struct DevState {
    bool ok = false;
    void *scratch = nullptr;
};
std::mutex g_init_mu;
std::map<int, DevState> g_devs;

int ensure_init()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_init_mu);
    DevState &ds = g_devs[dev];
    if (ds.ok) return BNTT_OK;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(BNTT_ENODEV, "device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    // this device's copy of the scratch array (device `dev` is current)
    void *sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!sp) return fail(BNTT_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
    ds.scratch = sp;
    ds.ok = true;
    return BNTT_OK;
}

// The current device's scratch address (after ensure_init on that device).
int noalloc_scratch(void **out)
{
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_init_mu);
    auto it = g_devs.find(dev);
    if (it == g_devs.end() || !it->second.ok || !it->second.scratch)
        return fail(BNTT_EHIP, "device %d is not initialised", dev);
    *out = it->second.scratch;
    return BNTT_OK;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Largest element count whose three arrays fit a size_t byte count.
constexpr size_t kMaxElems = SIZE_MAX / (3 * sizeof(bntt_u128));

// [p, p + bytes) and [q, q + bytes) share a byte but do not start together.
bool partial_overlap(const void *p, const void *q, size_t bytes)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p), b = reinterpret_cast<uintptr_t>(q);
    return a != b && a < b + bytes && b < a + bytes;
}

int check_vec(const void *c, const void *a, const void *b, size_t n, const bntt_u128 *m)
{
    if (!c || !a || !b || !m) return fail(BNTT_EINVAL, "null pointer argument");
    if (!aligned16(c) || !aligned16(a) || !aligned16(b))
        return fail(BNTT_EINVAL, "device arrays must be 16-byte aligned");
    if (n > kMaxElems) return fail(BNTT_EINVAL, "n = %zu is too large", n);
    // each element is read and written by one thread: c may be a or b, but
    // a shifted overlap would let one element's write race another's read
    const size_t bytes = n * sizeof(bntt_u128);
    if (partial_overlap(c, a, bytes) || partial_overlap(c, b, bytes))
        return fail(BNTT_EINVAL, "c must equal a (or b) or not overlap it");
    return BNTT_OK;
}

int blas_dev(int op, void *c, const void *a, const void *b, const bntt_u128 *x, size_t n, const bntt_u128 *m,
             const bntt_u128 *mu, void *stream, uint32_t *done = nullptr, uint32_t seq = 0)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n == 0) return BNTT_OK;
    rc = check_vec(c, a, b, n, m);
    if (rc) return rc;
    if ((op == bntt::OP_VVMUL || op == bntt::OP_AXPY) && !mu) return fail(BNTT_EINVAL, "mu is required");
    if (op == bntt::OP_AXPY && !x) return fail(BNTT_EINVAL, "axpy scalar is required");
    const Mod md = make_mod(m, mu);
    const U128 xs = x ? bntt::make_u128(x->lo, x->hi) : bntt::make_u128(0, 0);
    hipError_t e = bntt::launch_blas(op, static_cast<U128 *>(c), static_cast<const U128 *>(a),
                                     static_cast<const U128 *>(b), xs, n, md, static_cast<hipStream_t>(stream), done,
                                     seq);
    if (e != hipSuccess) return hip_fail(e, "blas kernel launch");
    return BNTT_OK;
}

// Library-owned staging for the host-pointer entry points: per device, one
// stream, one grow-only device arena and one grow-only pinned host buffer,
// kept until the process ends; calls on one device take turns.  Small calls
// run zero-copy: the caller's arrays are copied into the pinned buffer, the
// kernel reads and writes it over PCIe, and the result is copied out, with
// no DMA transfer.
struct Staging {
    std::mutex mu;
    hipStream_t stream = nullptr;
    void *buf = nullptr;
    size_t cap = 0;
    void *hbuf = nullptr;  // pinned host memory
    void *hdev = nullptr;  // its address for kernels
    size_t hcap = 0;
    uint32_t *flag_h = nullptr;  // pinned completion words (one per workgroup of a launch)
    uint32_t *flag_d = nullptr;  // their address for kernels
    uint32_t seq = 0;            // last sequence number posted to it
};
std::mutex g_stages_mu;
std::map<int, std::unique_ptr<Staging>> g_stages;

// The current device's staging (created empty at first use).
int stage_for_current(Staging **out)
{
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_stages_mu);
    std::unique_ptr<Staging> &st = g_stages[dev];
    if (!st) st.reset(new Staging());
    *out = st.get();
    return BNTT_OK;
}

// Completion words of the staging: one per workgroup of the widest launch that
// posts them (a BLAS op of 4096 residues, 32 workgroups; the split of a lone
// transform, 16).
constexpr size_t kFlagWords = 32;
static_assert(bntt::lone_parts(bntt::kLoneSplitMaxLog2) <= (int)kFlagWords, "a completion word per workgroup");

// Largest arrays (bytes each) that the host-pointer entry points move zero-copy.
constexpr size_t kZeroCopyBytes = (size_t)4096 * sizeof(bntt_u128);
// Host-array transforms up to this size (2^16 points) are staged through
// pinned memory, with DMA copies in and the result written straight back;
// larger ones copy the caller's (pageable) arrays with the runtime's own
// copies, which then run faster than one thread's copies into pinned memory.
constexpr size_t kStagedBytes = ((size_t)1 << 16) * sizeof(bntt_u128);

// The device arena, at least `bytes` (caller holds st.mu, current device = st's).
int stage_reserve(Staging &stg, size_t bytes)
{
    hipError_t e;
    if (!stg.stream) {
        e = hipStreamCreate(&stg.stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    }
    if (stg.cap < bytes) {
        if (stg.buf) {
            e = hipStreamSynchronize(stg.stream);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            (void)hipFree(stg.buf);
            stg.buf = nullptr;
            stg.cap = 0;
        }
        e = hipMalloc(&stg.buf, bytes);
        if (e != hipSuccess) return fail(BNTT_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        stg.cap = bytes;
    }
    return BNTT_OK;
}

// The pinned completion words of the staging (allocated at first use).
int stage_flag(Staging &stg)
{
    if (stg.flag_h) return BNTT_OK;
    void *h = nullptr, *d = nullptr;
    hipError_t e = hipHostMalloc(&h, kFlagWords * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(BNTT_ENOMEM, "hipHostMalloc(%zu): %s", kFlagWords * sizeof(uint32_t), hipGetErrorString(e));
    e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        return hip_fail(e, "hipHostGetDevicePointer");
    }
    stg.flag_h = static_cast<uint32_t *>(h);
    stg.flag_d = static_cast<uint32_t *>(d);
    // one word per workgroup of a launch that posts completion words
    for (size_t w = 0; w < kFlagWords; w++) __atomic_store_n(stg.flag_h + w, stg.seq, __ATOMIC_RELEASE);
    return BNTT_OK;
}

// Waits for a launch on the staging's stream that posts `seq` to its
// completion word.  The host sees that store several microseconds before the
// stream's own completion, so it spins on the word, for at most about 2 ms;
// only if the word has not arrived by then is the stream synchronised (which
// also reports a failed launch).
int wait_flag(Staging &stg, uint32_t seq, int words = 1)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    int w = 0;  // words [0, w) have arrived
    for (uint32_t i = 1;; i++) {
        while (w < words && __atomic_load_n(stg.flag_h + w, __ATOMIC_ACQUIRE) == seq) w++;
        if (w == words) return BNTT_OK;
        if ((i & 255u) == 0 && clk::now() - t0 > std::chrono::milliseconds(2)) break;
    }
    const hipError_t e = hipStreamSynchronize(stg.stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    for (; w < words; w++) {
        const uint32_t v = __atomic_load_n(stg.flag_h + w, __ATOMIC_ACQUIRE);
        if (v == (seq ^ 0x80000000u))
            return fail(BNTT_EHIP, "workgroup %d of the transform gave up waiting for its siblings at the grid "
                                   "barrier (the GPU did not run them within about a second); results discarded", w);
        if (v != seq) return fail(BNTT_EHIP, "the transform completed without posting its completion word");
    }
    return BNTT_OK;
}

// The pinned host buffer, at least `bytes` (after stage_reserve).
int stage_host_reserve(Staging &stg, size_t bytes)
{
    if (stg.hcap >= bytes) return BNTT_OK;
    hipError_t e;
    if (stg.hbuf) {
        e = hipStreamSynchronize(stg.stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        (void)hipHostFree(stg.hbuf);
        stg.hbuf = stg.hdev = nullptr;
        stg.hcap = 0;
    }
    e = hipHostMalloc(&stg.hbuf, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(BNTT_ENOMEM, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipHostGetDevicePointer(&stg.hdev, stg.hbuf, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(stg.hbuf);
        stg.hbuf = nullptr;
        return hip_fail(e, "hipHostGetDevicePointer");
    }
    stg.hcap = bytes;
    return BNTT_OK;
}

int blas_host(int op, bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, const bntt_u128 *x, size_t n,
              const bntt_u128 *m, const bntt_u128 *mu)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n == 0) return BNTT_OK;
    if (!c || !a || !b || !m) return fail(BNTT_EINVAL, "null pointer argument");
    if (n > kMaxElems) return fail(BNTT_EINVAL, "n = %zu is too large", n);
    const size_t bytes = n * sizeof(bntt_u128);
    Staging *st = nullptr;
    if ((rc = stage_for_current(&st))) return rc;
    Staging &stg = *st;
    std::lock_guard<std::mutex> lk(stg.mu);
    rc = stage_reserve(stg, 3 * bytes);
    if (rc) return rc;
    hipStream_t s = stg.stream;
    if (bytes <= kZeroCopyBytes) {
        // the kernel reads a and b from pinned memory and writes c there;
        // each of its workgroups posts a completion word, which the call
        // waits on instead of the stream
        if ((rc = stage_host_reserve(stg, 3 * bytes)) || (rc = stage_flag(stg))) return rc;
        char *hb = static_cast<char *>(stg.hbuf), *hd = static_cast<char *>(stg.hdev);
        std::memcpy(hb, a, bytes);
        std::memcpy(hb + bytes, b, bytes);
        const size_t words = bntt::blas_blocks(n);
        if (words > kFlagWords) return fail(BNTT_EHIP, "internal: %zu completion words needed", words);
        const uint32_t seq = ++stg.seq;
        rc = blas_dev(op, hd + 2 * bytes, hd, hd + bytes, x, n, m, mu, s, stg.flag_d, seq);
        if (rc) return rc;
        if ((rc = wait_flag(stg, seq, (int)words))) return rc;
        std::memcpy(c, hb + 2 * bytes, bytes);
        return BNTT_OK;
    }
    char *base = static_cast<char *>(stg.buf);
    void *da = base, *db = base + bytes, *dc = base + 2 * bytes;
    hipError_t e = hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
    rc = blas_dev(op, dc, da, db, x, n, m, mu, s);
    if (rc) return rc;
    e = hipMemcpyAsync(c, dc, bytes, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return BNTT_OK;
}

}  // namespace

int bntt_set_error(int code, const char *msg) { return fail(code, "%s", msg); }
int bntt_set_hip_error(hipError_t e, const char *what) { return hip_fail(e, what); }
int bntt_ensure_init() { return ensure_init(); }
bntt::Mod bntt_make_mod(const bntt_u128 *m, const bntt_u128 *mu) { return make_mod(m, mu); }

#pragma GCC visibility push(default)
extern "C" {

int bntt_version(void) { return 100; }

int bntt_init(void) { return ensure_init(); }

const char *bntt_last_error(void) { return g_err.c_str(); }

int bntt_vvadd_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, void *stream)
{
    return blas_dev(bntt::OP_VVADD, c, a, b, nullptr, n, m, nullptr, stream);
}

int bntt_vvsub_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, void *stream)
{
    return blas_dev(bntt::OP_VVSUB, c, a, b, nullptr, n, m, nullptr, stream);
}

int bntt_vvmul_dev(void *c, const void *a, const void *b, size_t n, const bntt_u128 *m, const bntt_u128 *mu,
                   void *stream)
{
    return blas_dev(bntt::OP_VVMUL, c, a, b, nullptr, n, m, mu, stream);
}

int bntt_axpy_dev(void *c, const void *a, const void *b, const bntt_u128 *x, size_t n, const bntt_u128 *m,
                  const bntt_u128 *mu, void *stream)
{
    return blas_dev(bntt::OP_AXPY, c, a, b, x, n, m, mu, stream);
}

size_t bntt_ntt_tmp_bytes(int log2n, size_t count)
{
    if (log2n < bntt::kMinLog2 || log2n > bntt::kMaxLog2) return 0;
    // one pass on one workgroup per transform, except a lone transform of
    // 2^11 .. 2^13 points, which with a tmp (its intermediate and 16 residues
    // of grid-barrier words) splits over several workgroups; above 4096 points
    // the tmp is required (the intermediate of several passes)
    if (count == 1 && log2n >= bntt::kLoneSplitMinLog2 && log2n <= bntt::kLoneSplitMaxLog2)
        return (((size_t)1 << log2n) + 16) * sizeof(bntt_u128);
    if (log2n <= bntt::kLdsMaxLog2) return 0;
    if (count > (SIZE_MAX / sizeof(bntt_u128)) >> log2n) return SIZE_MAX;  // no such buffer exists
    return (count << log2n) * sizeof(bntt_u128);
}

int bntt_ntt_fwd_dev(void *y, const void *x, const void *tw, int log2n, size_t count, const bntt_u128 *m,
                     const bntt_u128 *mu, void *tmp, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (log2n < bntt::kMinLog2 || log2n > bntt::kMaxLog2)
        return fail(BNTT_EINVAL, "log2n = %d outside [%d, %d]", log2n, bntt::kMinLog2, bntt::kMaxLog2);
    if (count == 0) return BNTT_OK;
    if (!y || !x || !tw || !m || !mu) return fail(BNTT_EINVAL, "null pointer argument");
    if (!aligned16(y) || !aligned16(x) || !aligned16(tw) || (tmp && !aligned16(tmp)))
        return fail(BNTT_EINVAL, "device arrays must be 16-byte aligned");
    if (count > (SIZE_MAX / sizeof(bntt_u128)) >> log2n) return fail(BNTT_EINVAL, "count = %zu is too large", count);
    const size_t n = (size_t)1 << log2n;
    const size_t bytes = count * n * sizeof(bntt_u128);
    if (log2n > bntt::kLdsMaxLog2) {
        if (!tmp) return fail(BNTT_EINVAL, "log2n = %d needs a tmp buffer of bntt_ntt_tmp_bytes()", log2n);
        const char *yb = static_cast<const char *>(y), *xb = static_cast<const char *>(x);
        if (yb < xb + bytes && xb < yb + bytes) return fail(BNTT_EINVAL, "y must not overlap x for log2n > 12");
    } else if (partial_overlap(y, x, bytes)) {
        // one workgroup's stores would land in another transform's input
        // before that transform is loaded
        return fail(BNTT_EINVAL, "y must equal x or not overlap it");
    }
    const size_t tmpb = bntt_ntt_tmp_bytes(log2n, count);
    if (tmpb == 0) {
        tmp = nullptr;  // one pass: a tmp given anyway is not touched
    } else if (tmp) {
        const char *tb = static_cast<const char *>(tmp), *yb = static_cast<const char *>(y),
                   *xb = static_cast<const char *>(x);
        if ((tb < xb + bytes && xb < tb + tmpb) || (tb < yb + bytes && yb < tb + tmpb))
            return fail(BNTT_EINVAL, "tmp must not overlap x or y");
    }
    const Mod md = make_mod(m, mu);
    hipError_t e = bntt::launch_ntt_uniform(static_cast<U128 *>(y), static_cast<const U128 *>(x),
                                            static_cast<const U128 *>(tw), log2n, count, md,
                                            static_cast<U128 *>(tmp), static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "ntt kernel launch");
    return BNTT_OK;
}

int bntt_ntt_fwd_dev_noalloc(void *y, const void *x, const void *tw, int log2n, size_t count, const bntt_u128 *m,
                             const bntt_u128 *mu, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (log2n < bntt::kMinLog2 || log2n > bntt::kMaxLog2)
        return fail(BNTT_EINVAL, "log2n = %d outside [%d, %d]", log2n, bntt::kMinLog2, bntt::kMaxLog2);
    void *tmp = nullptr;
    if (count > 0 && bntt_ntt_tmp_bytes(log2n, count) != 0) {
        // multi-pass sizes, and a lone transform of 2^11 .. 2^13 points split
        // over several workgroups, keep their intermediates in the fixed
        // scratch array; every element of it that a later pass reads is written
        // by an earlier pass of the same call, and the split's barrier words
        // are matched against a token no earlier launch used, so nothing needs
        // clearing
        if (count > (kNoallocResidues >> log2n))
            return fail(BNTT_EINVAL,
                        "count * 2^log2n = %zu * 2^%d exceeds the %zu residues of the no-allocation scratch "
                        "(use bntt_ntt_fwd_dev with a tmp buffer)",
                        count, log2n, kNoallocResidues);
        if ((rc = noalloc_scratch(&tmp))) return rc;
    }
    return bntt_ntt_fwd_dev(y, x, tw, log2n, count, m, mu, tmp, stream);
}

int bntt_vvadd(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m)
{
    return blas_host(bntt::OP_VVADD, c, a, b, nullptr, n, m, nullptr);
}

int bntt_vvsub(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m)
{
    return blas_host(bntt::OP_VVSUB, c, a, b, nullptr, n, m, nullptr);
}

int bntt_vvmul(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, size_t n, const bntt_u128 *m,
               const bntt_u128 *mu)
{
    return blas_host(bntt::OP_VVMUL, c, a, b, nullptr, n, m, mu);
}

int bntt_axpy(bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, const bntt_u128 *x, size_t n,
              const bntt_u128 *m, const bntt_u128 *mu)
{
    return blas_host(bntt::OP_AXPY, c, a, b, x, n, m, mu);
}

int bntt_ntt_fwd(bntt_u128 *Y, const bntt_u128 *X, const bntt_u128 *tw, int log2n, const bntt_u128 *m,
                 const bntt_u128 *mu)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (log2n < bntt::kMinLog2 || log2n > bntt::kMaxLog2)
        return fail(BNTT_EINVAL, "log2n = %d outside [%d, %d]", log2n, bntt::kMinLog2, bntt::kMaxLog2);
    if (!Y || !X || !tw || !m || !mu) return fail(BNTT_EINVAL, "null pointer argument");
    const size_t n = (size_t)1 << log2n;
    const size_t bytes = n * sizeof(bntt_u128);
    const size_t tmpb = bntt_ntt_tmp_bytes(log2n, 1);
    Staging *st = nullptr;
    if ((rc = stage_for_current(&st))) return rc;
    Staging &stg = *st;
    std::lock_guard<std::mutex> lk(stg.mu);
    rc = stage_reserve(stg, 3 * bytes + tmpb);
    if (rc) return rc;
    hipStream_t s = stg.stream;
    if (bytes <= kZeroCopyBytes) {
        // one launch: the kernel reads x and the table once from pinned
        // memory, writes y there once and each of its workgroups posts a
        // completion word (a transform of 2^11 or 2^12 points splits over
        // several workgroups, with its intermediate in the device arena)
        if ((rc = stage_host_reserve(stg, 3 * bytes)) || (rc = stage_flag(stg))) return rc;
        char *hb = static_cast<char *>(stg.hbuf), *hd = static_cast<char *>(stg.hdev);
        std::memcpy(hb, X, bytes);
        std::memcpy(hb + bytes, tw, bytes);
        const uint32_t seq = ++stg.seq;
        int words = 1;
        const hipError_t e =
            bntt::launch_ntt_uniform(reinterpret_cast<U128 *>(hd + 2 * bytes), reinterpret_cast<const U128 *>(hd),
                                     reinterpret_cast<const U128 *>(hd + bytes), log2n, 1, make_mod(m, mu),
                                     tmpb ? static_cast<U128 *>(stg.buf) : nullptr, s, stg.flag_d, seq, &words,
                                     (int)kFlagWords);
        if (e != hipSuccess) return hip_fail(e, "ntt kernel launch");
        if ((rc = wait_flag(stg, seq, words))) return rc;
        std::memcpy(Y, hb + 2 * bytes, bytes);
        return BNTT_OK;
    }
    char *base = static_cast<char *>(stg.buf);
    void *dx = base, *dtw = base + bytes, *dy = base + 2 * bytes;
    void *dtmp = tmpb ? base + 3 * bytes : nullptr;
    if (bytes <= kStagedBytes) {
        // the caller's table is copied into pinned memory and DMA moves it to
        // the device while x is copied into pinned memory; the transform reads
        // x there (each element once) and the table in device memory, writes y
        // straight into pinned memory, and y is copied out once the launch's
        // completion words (one launch over several workgroups) or the stream
        // (several passes) say it is there
        if ((rc = stage_host_reserve(stg, 3 * bytes)) || (rc = stage_flag(stg))) return rc;
        char *hb = static_cast<char *>(stg.hbuf), *hd = static_cast<char *>(stg.hdev);
        std::memcpy(hb + bytes, tw, bytes);
        hipError_t e = hipMemcpyAsync(dtw, hb + bytes, bytes, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
        std::memcpy(hb, X, bytes);
        dx = hd;
        U128 *yh = reinterpret_cast<U128 *>(hd + 2 * bytes);
        if (log2n <= bntt::kLoneSplitMaxLog2) {
            const uint32_t seq = ++stg.seq;
            int words = 1;
            e = bntt::launch_ntt_uniform(yh, static_cast<const U128 *>(dx), static_cast<const U128 *>(dtw), log2n, 1,
                                         make_mod(m, mu), static_cast<U128 *>(dtmp), s, stg.flag_d, seq, &words,
                                         (int)kFlagWords);
            if (e != hipSuccess) return hip_fail(e, "ntt kernel launch");
            if ((rc = wait_flag(stg, seq, words))) return rc;
        } else {
            // several passes: completion words from the last pass's 32..128
            // workgroups, each releasing at system scope, measured slower
            // than the stream (58.9 against 53.4 us at 2^14)
            if ((rc = bntt_ntt_fwd_dev(yh, dx, dtw, log2n, 1, m, mu, dtmp, s))) return rc;
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        }
        std::memcpy(Y, hb + 2 * bytes, bytes);
        return BNTT_OK;
    }
    hipError_t e = hipMemcpyAsync(dx, X, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dtw, tw, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
    rc = bntt_ntt_fwd_dev(dy, dx, dtw, log2n, 1, m, mu, dtmp, s);
    if (rc) return rc;
    e = hipMemcpyAsync(Y, dy, bytes, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return BNTT_OK;
}

}  // extern "C"
#pragma GCC visibility pop
