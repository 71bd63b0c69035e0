// u128.hpp -- 128-bit modular arithmetic for gfx950 (also compiles for the host).
//
// A residue is four 32-bit limbs, least significant first: the native memory
// layout of unsigned __int128, so one element is one 16-byte vector load
// (global_load_dwordx4 / ds_read_b128).  Products are built from 32x32->64
// multiply-adds, which the compiler maps to v_mad_u64_u32; carries use the
// add/sub-with-carry builtins, which map to v_add_co/v_addc/v_sub_co/v_subb.
//
// Semantics are those of the reference's scalar primitives, bit for bit and
// for every 128-bit input (citations are to the reference's src/):
//   addmod_ref   ntt_scalar.c:49-79   literal, including its edge behaviour
//                (when lo64(a+b) == lo64(m) and a+b >= m the high word borrows
//                from !c5, so the result is a+b-m-2^64, or m when a+b == m)
//   submod_ref   ntt_scalar.c:81-104  literal
//   mulmod_ref   ntt_scalar.c:106-199 literal Barrett with the MBITS=124 shift
//                schedule, including the 128-bit wraps of mul128/mulhi128
//   mulmod_pm    exact a*b mod q for q = 2^124 - c, c < 2^16, a,b < 2^124.
//                When mu = floor(2^251/q) and q is in (2^123, 2^124), the
//                reference's Barrett quotient is short by at most one, so its
//                single conditional subtract returns exactly a*b mod q for all
//                a, b < 2^124; mulmod_pm returns the same value with about half
//                the multiplies (two folds of 2^124 = c mod q).
//   mulmod_r31   the same exact product with a in 31-bit digits (the second
//                operand pre-converted: twiddles, the axpy scalar): 16
//                multiply-adds with no carry fix-ups; the fast path of the
//                network and of vvmul/axpy.
//   mulmod       mulmod_pm when the modulus qualifies and both operands are
//                below 2^124, mulmod_ref otherwise: bit-exact in every case.
// and, for the polynomial-multiply path (any odd q < 2^122, exact results):
//   montmul_r31_lazy / montmul_r31   Montgomery product a*w*2^-124, below 2q
//                / canonical, for a < 4q and w < q
//   addmod_canon / submod_canon      canonical add/sub; add128l / sub128l /
//                sub_if_ge for lazy butterflies (values below 4q).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bntt {

#define BNTT_HD __host__ __device__ __forceinline__

struct alignas(16) U128 {
    uint32_t v[4];
};

// A residue as four 31-bit digits: value = d0 + d1*2^31 + d2*2^62 + d3*2^93.
struct R31 {
    uint32_t d[4];
};

// Montgomery constants of an odd modulus q < 2^122 (R = 2^124).
struct MontQ {
    R31 q;       // q in digits
    R31 qinv;    // -q^-1 mod 2^124 in digits
    U128 ql;     // q in 32-bit limbs
    U128 ql2;    // 2q in 32-bit limbs
};

// Modulus descriptor passed to kernels by value.
struct Mod {
    U128 m;       // modulus
    U128 mu;      // Barrett constant of the reference formula
    U128 notm;    // ~m (used by the literal addmod)
    uint32_t c;   // 2^124 - m when pm != 0
    uint32_t pm;  // m = 2^124 - c, c < 2^16, mu = floor(2^251/m)
    MontQ mq;     // exact Montgomery constants (polynomial path only)
};

BNTT_HD U128 make_u128(uint64_t lo, uint64_t hi)
{
    U128 r;
    r.v[0] = (uint32_t)lo;
    r.v[1] = (uint32_t)(lo >> 32);
    r.v[2] = (uint32_t)hi;
    r.v[3] = (uint32_t)(hi >> 32);
    return r;
}

BNTT_HD uint64_t lo64(const U128 &a) { return (uint64_t)a.v[0] | ((uint64_t)a.v[1] << 32); }
BNTT_HD uint64_t hi64(const U128 &a) { return (uint64_t)a.v[2] | ((uint64_t)a.v[3] << 32); }

// 32-bit add/sub with carry (v_add_co_u32 / v_addc_co_u32 / v_sub_co_u32 / v_subb_co_u32).
BNTT_HD uint32_t addc(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout)
{
    unsigned int co;
    uint32_t r = __builtin_addc(a, b, cin, &co);
    cout = co;
    return r;
}
BNTT_HD uint32_t subb(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout)
{
    unsigned int bo;
    uint32_t r = __builtin_subc(a, b, bin, &bo);
    bout = bo;
    return r;
}

// (hi:lo) >> s for 0 < s < 32 (v_alignbit_b32).
BNTT_HD uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
}

BNTT_HD uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

// The predicate holds in some active lane of the wave (on the host: in the one
// lane there is).  Lets a wave skip a rarely needed exact step as a whole.
BNTT_HD bool any_lane(bool p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(p) != 0;
#else
    return p;
#endif
}

// ---------------------------------------------------------------- literal add/sub

// ntt_scalar.c:49-79, literal.  notm = ~m.
BNTT_HD U128 addmod_ref(const U128 &a, const U128 &b, const U128 &m, const U128 &notm)
{
    uint32_t k, s0, s1, s2, s3;
    s0 = addc(a.v[0], b.v[0], 0, k);
    s1 = addc(a.v[1], b.v[1], k, k);
    s2 = addc(a.v[2], b.v[2], k, k);
    s3 = addc(a.v[3], b.v[3], k, k);
    // c5 = borrow of (ml - t18), i.e. t18 > ml
    uint32_t c5;
    (void)subb(m.v[0], s0, 0, c5);
    (void)subb(m.v[1], s1, c5, c5);
    // d4 = t18 - ml (mod 2^64)
    uint32_t bb, d0, d1, d2, d3;
    d0 = subb(s0, m.v[0], 0, bb);
    d1 = subb(s1, m.v[1], bb, bb);
    // d6 = t19 - mh - !c5 = t19 + ~mh + c5 (mod 2^64); the carry out of this
    // sum is i14, the borrow of (mh - t19 - c5) over 128 bits: t19 + ~mh + c5
    // reaches 2^64 exactly when t19 + c5 > mh
    uint32_t i14;
    d2 = addc(s2, notm.v[2], c5, bb);
    d3 = addc(s3, notm.v[3], bb, i14);
    U128 r;
    r.v[0] = i14 ? d0 : s0;
    r.v[1] = i14 ? d1 : s1;
    r.v[2] = i14 ? d2 : s2;
    r.v[3] = i14 ? d3 : s3;
    return r;
}

// ntt_scalar.c:81-104, literal: a - b, plus m on a borrow out of 128 bits.
BNTT_HD U128 submod_ref(const U128 &a, const U128 &b, const U128 &m)
{
    uint32_t k, d0, d1, d2, d3, e0, e1, e2, e3, kk;
    d0 = subb(a.v[0], b.v[0], 0, k);
    d1 = subb(a.v[1], b.v[1], k, k);
    d2 = subb(a.v[2], b.v[2], k, k);
    d3 = subb(a.v[3], b.v[3], k, k);
    e0 = addc(d0, m.v[0], 0, kk);
    e1 = addc(d1, m.v[1], kk, kk);
    e2 = addc(d2, m.v[2], kk, kk);
    e3 = addc(d3, m.v[3], kk, kk);
    U128 r;
    r.v[0] = k ? e0 : d0;
    r.v[1] = k ? e1 : d1;
    r.v[2] = k ? e2 : d2;
    r.v[3] = k ? e3 : d3;
    return r;
}

// The same two formulas for a modulus of the fast form m = 2^124 - c (c < 2^16):
// limbs 1..3 of m are 0xFFFFFFFF, 0xFFFFFFFF, 0x0FFFFFFF and those of ~m 0, 0,
// 0xF0000000 for every such m, so they enter as constants (the first two as
// inline operands, which an add with carry-in can take beside its carry,
// where a register operand would first be copied into a vector register).
// Same operations, same bits, for all inputs.
BNTT_HD U128 addmod_pm(const U128 &a, const U128 &b, uint32_t m0)
{
    U128 m, notm;
    m.v[0] = m0;
    m.v[1] = 0xFFFFFFFFu;
    m.v[2] = 0xFFFFFFFFu;
    m.v[3] = 0x0FFFFFFFu;
    notm.v[0] = ~m0;
    notm.v[1] = 0u;
    notm.v[2] = 0u;
    notm.v[3] = 0xF0000000u;
    return addmod_ref(a, b, m, notm);
}
BNTT_HD U128 submod_pm(const U128 &a, const U128 &b, uint32_t m0)
{
    U128 m;
    m.v[0] = m0;
    m.v[1] = 0xFFFFFFFFu;
    m.v[2] = 0xFFFFFFFFu;
    m.v[3] = 0x0FFFFFFFu;
    return submod_ref(a, b, m);
}

// ------------------------------------------------------------------- products

// Full 256-bit product of two 128-bit values (schoolbook, 16 multiply-adds).
// Exact for all inputs.
BNTT_HD void mul_4x4(const U128 &a, const U128 &b, uint32_t p[8])
{
    uint64_t t;
    t = (uint64_t)a.v[0] * b.v[0];
    p[0] = (uint32_t)t;
    t = mad32(a.v[0], b.v[1], t >> 32);
    p[1] = (uint32_t)t;
    t = mad32(a.v[0], b.v[2], t >> 32);
    p[2] = (uint32_t)t;
    t = mad32(a.v[0], b.v[3], t >> 32);
    p[3] = (uint32_t)t;
    p[4] = (uint32_t)(t >> 32);
#pragma unroll
    for (int i = 1; i < 4; i++) {
        t = mad32(a.v[i], b.v[0], p[i]);
        p[i] = (uint32_t)t;
        t = mad32(a.v[i], b.v[1], (uint64_t)p[i + 1] + (t >> 32));
        p[i + 1] = (uint32_t)t;
        t = mad32(a.v[i], b.v[2], (uint64_t)p[i + 2] + (t >> 32));
        p[i + 2] = (uint32_t)t;
        t = mad32(a.v[i], b.v[3], (uint64_t)p[i + 3] + (t >> 32));
        p[i + 3] = (uint32_t)t;
        p[i + 4] = (uint32_t)(t >> 32);
    }
}

// ---------------------------------------------------------- fast exact mulmod

// a*b mod q for q = 2^124 - c, c < 2^16, a, b < 2^124.  Canonical result.
//   x = a*b < 2^248 = xh*2^124 + xl
//   t = xl + c*xh            < 2^124 + 2^140, t >> 124 <= c
//   u = (t mod 2^124) + c*(t >> 124)  < 2^124 + 2^32 < 2q
//   r = u >= q ? u - q : u, with u >= q  <=>  u + c >= 2^124.
BNTT_HD U128 mulmod_pm(const U128 &a, const U128 &b, uint32_t c)
{
    uint32_t p[8];
    mul_4x4(a, b, p);
    const uint32_t h0 = funnel(p[4], p[3], 28), h1 = funnel(p[5], p[4], 28);
    const uint32_t h2 = funnel(p[6], p[5], 28), h3 = funnel(p[7], p[6], 28);
    uint64_t t = mad32(c, h0, p[0]);
    const uint32_t t0 = (uint32_t)t;
    t = mad32(c, h1, (uint64_t)p[1] + (t >> 32));
    const uint32_t t1 = (uint32_t)t;
    t = mad32(c, h2, (uint64_t)p[2] + (t >> 32));
    const uint32_t t2 = (uint32_t)t;
    t = mad32(c, h3, (uint64_t)(p[3] & 0x0FFFFFFFu) + (t >> 32));
    const uint32_t th = (uint32_t)(t >> 28);  // <= c < 2^16
    const uint32_t t3 = (uint32_t)t & 0x0FFFFFFFu;
    uint32_t k, u0, u1, u2, u3;
    u0 = addc(t0, c * th, 0, k);  // c*th < 2^32
    u1 = addc(t1, 0, k, k);
    u2 = addc(t2, 0, k, k);
    u3 = t3 + k;  // < 2^28 + 1
    uint32_t d0, d1, d2, d3;
    d0 = addc(u0, c, 0, k);
    d1 = addc(u1, 0, k, k);
    d2 = addc(u2, 0, k, k);
    d3 = u3 + k;
    const bool ge = (d3 >> 28) != 0;
    U128 r;
    r.v[0] = ge ? d0 : u0;
    r.v[1] = ge ? d1 : u1;
    r.v[2] = ge ? d2 : u2;
    r.v[3] = ge ? (d3 & 0x0FFFFFFFu) : u3;
    return r;
}

// ------------------------------------------------- radix-2^31 exact mulmod
//
// On gfx950 v_mad_u64_u32 (32x32 -> 64 plus a 64-bit addend) issues at the
// same rate as a 32-bit add, so the cost of a 128-bit product is its
// instruction count, and in radix 2^32 most of that is carry handling.  With
// the residue split into four 31-bit digits (124 = 4 * 31) every partial
// product is below 2^62 and a column of four of them plus the incoming carry
// stays below 2^64, so each column is a plain chain of multiply-adds with no
// carry fixups: 16 v_mad_u64_u32 plus one mask and one 64-bit shift per column.

constexpr uint32_t kM31 = 0x7FFFFFFFu;

// 32-bit limbs -> 31-bit digits, for a < 2^124.
BNTT_HD R31 to_r31(const U128 &a)
{
    R31 r;
    r.d[0] = a.v[0] & kM31;
    r.d[1] = funnel(a.v[1], a.v[0], 31) & kM31;
    r.d[2] = funnel(a.v[2], a.v[1], 30) & kM31;
    r.d[3] = funnel(a.v[3], a.v[2], 29);
    return r;
}

// a*w mod q for q = 2^124 - c (c < 2^16) and a, w < 2^124; w given in digits.
// Canonical result in 32-bit limbs.
//   x = a*w = sum_k d_k 2^(31k), k < 8      (columns normalized on the fly)
//   V = sum_{k<4} (d_k + c*d_{k+4}) 2^(31k) = f + t*2^124,  t <= c
//   W = f + c*t < 2^124 + 2^32,  result = W >= q ? W - q : W,
//   with W >= q  <=>  W + c >= 2^124.
BNTT_HD U128 mulmod_r31(const U128 &a_, const R31 &w, uint32_t c)
{
    const R31 a = to_r31(a_);
    uint64_t t;
    uint32_t d[8];
    t = (uint64_t)a.d[0] * w.d[0];
    d[0] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[1], t >> 31);
    t = mad32(a.d[1], w.d[0], t);
    d[1] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[2], t >> 31);
    t = mad32(a.d[1], w.d[1], t);
    t = mad32(a.d[2], w.d[0], t);
    d[2] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[3], t >> 31);
    t = mad32(a.d[1], w.d[2], t);
    t = mad32(a.d[2], w.d[1], t);
    t = mad32(a.d[3], w.d[0], t);
    d[3] = (uint32_t)t & kM31;
    t = mad32(a.d[1], w.d[3], t >> 31);
    t = mad32(a.d[2], w.d[2], t);
    t = mad32(a.d[3], w.d[1], t);
    d[4] = (uint32_t)t & kM31;
    t = mad32(a.d[2], w.d[3], t >> 31);
    t = mad32(a.d[3], w.d[2], t);
    d[5] = (uint32_t)t & kM31;
    t = mad32(a.d[3], w.d[3], t >> 31);
    d[6] = (uint32_t)t & kM31;
    d[7] = (uint32_t)(t >> 31);
    // fold 2^124 = c (mod q): digits k and k+4 together, carries below 2^18
    uint32_t f[4];
    t = mad32(c, d[4], d[0]);
    f[0] = (uint32_t)t & kM31;
    t = mad32(c, d[5], (t >> 31) + d[1]);
    f[1] = (uint32_t)t & kM31;
    t = mad32(c, d[6], (t >> 31) + d[2]);
    f[2] = (uint32_t)t & kM31;
    t = mad32(c, d[7], (t >> 31) + d[3]);
    f[3] = (uint32_t)t & kM31;
    const uint32_t top = (uint32_t)(t >> 31);  // t <= c
    // back to 32-bit limbs (f < 2^124), then add c*top
    uint32_t k, w0, w1, w2, w3;
    w0 = addc(f[0] | (f[1] << 31), c * top, 0, k);
    w1 = addc((f[1] >> 1) | (f[2] << 30), 0, k, k);
    w2 = addc((f[2] >> 2) | (f[3] << 29), 0, k, k);
    w3 = (f[3] >> 3) + k;  // <= 2^28
    // w < 2^124 + 2^20 is congruent to a*w; it reaches q = 2^124 - c only with
    // a top limb of 0x0FFFFFFF or more (about one product in 2^28 on uniform
    // operands), so a wave skips the exact correction below unless one of its
    // lanes may need it, and then every lane takes it
    if (!any_lane(w3 >= 0x0FFFFFFFu)) {
        U128 r;
        r.v[0] = w0;
        r.v[1] = w1;
        r.v[2] = w2;
        r.v[3] = w3;
        return r;
    }
    uint32_t z0, z1, z2, z3;
    z0 = addc(w0, c, 0, k);
    z1 = addc(w1, 0, k, k);
    z2 = addc(w2, 0, k, k);
    z3 = w3 + k;
    const bool ge = (z3 >> 28) != 0;
    U128 r;
    r.v[0] = ge ? z0 : w0;
    r.v[1] = ge ? z1 : w1;
    r.v[2] = ge ? z2 : w2;
    r.v[3] = ge ? (z3 & 0x0FFFFFFFu) : w3;
    return r;
}

// -------------------------------------- exact Montgomery product, any odd q < 2^122
//
// For the polynomial-multiply path (no counterpart in the reference), whose
// modulus q_116 = 2^115 + 47*2^17 + 1 is not of the 2^124 - c form and for which
// the reference's MBITS=124 Barrett schedule is not exact.  R = 2^124 (four
// digits), qinv = -q^-1 mod R.  For a, w < q the result a*w*R^-1 mod q is
// canonical: u = a*w + m*q with m = (a*w mod R)*qinv mod R is divisible by R
// and u/R < (q^2 + R*q)/R < 2q, so one conditional subtraction finishes.

// Montgomery product a * w * 2^-124 mod q for odd q < 2^122, without the final
// subtraction: for a < 4q and w < q the result is below 2q (T + m q < 4q^2 +
// 2^124 q, and 4q < 2^124), which lazy butterflies keep as is.  SPARSE: q's
// digits 1 and 2 are zero (q = d0 + d3 2^93, such as 2^115 + 47 2^17 + 1), so
// m*q takes 8 multiply-adds instead of 16.
template <bool SPARSE = false>
BNTT_HD U128 montmul_r31_lazy(const U128 &a_, const R31 &w, const MontQ &mq)
{
    const R31 a = to_r31(a_);
    uint64_t t;
    if constexpr (SPARSE) {
        // q's digits 1 and 2 are zero, so a column holds at most four digit
        // products of a*w and two of m*q: the reduction runs inside the
        // product, column by column, every sum a 64-bit multiply-add on the
        // running column (no digit extracted between the product and the
        // reduction).  Bound, for q < 2^121, a < 4q, w < q, D = 2^31 - 1: the
        // fullest column (digit 3) is below 3 D^2 + 6 D 2^28 + 2^34 < 2^64.
        // m_i clears digit i of the running sum, as below; m is the same
        // unique m, so are the bits.  The caller selects SPARSE only for
        // q < 2^121.
        const R31 &q = mq.q;
        const uint32_t qi0 = mq.qinv.d[0];
        uint32_t m0, m1, m2, m3;
        t = (uint64_t)a.d[0] * w.d[0];
        m0 = ((uint32_t)t * qi0) & kM31;
        t = mad32(m0, q.d[0], t);
        t = mad32(a.d[0], w.d[1], t >> 31);
        t = mad32(a.d[1], w.d[0], t);
        m1 = ((uint32_t)t * qi0) & kM31;
        t = mad32(m1, q.d[0], t);
        t = mad32(a.d[0], w.d[2], t >> 31);
        t = mad32(a.d[1], w.d[1], t);
        t = mad32(a.d[2], w.d[0], t);
        m2 = ((uint32_t)t * qi0) & kM31;
        t = mad32(m2, q.d[0], t);
        t = mad32(a.d[0], w.d[3], t >> 31);
        t = mad32(a.d[1], w.d[2], t);
        t = mad32(a.d[2], w.d[1], t);
        t = mad32(a.d[3], w.d[0], t);
        t = mad32(m0, q.d[3], t);
        m3 = ((uint32_t)t * qi0) & kM31;
        t = mad32(m3, q.d[0], t);
        t = mad32(a.d[1], w.d[3], t >> 31);
        t = mad32(a.d[2], w.d[2], t);
        t = mad32(a.d[3], w.d[1], t);
        t = mad32(m1, q.d[3], t);
        const uint32_t r0 = (uint32_t)t & kM31;
        t = mad32(a.d[2], w.d[3], t >> 31);
        t = mad32(a.d[3], w.d[2], t);
        t = mad32(m2, q.d[3], t);
        const uint32_t r1 = (uint32_t)t & kM31;
        t = mad32(a.d[3], w.d[3], t >> 31);
        t = mad32(m3, q.d[3], t);
        const uint32_t r2 = (uint32_t)t & kM31;
        const uint32_t r3 = (uint32_t)(t >> 31);  // u/R < 2q < 2^123
        U128 u;
        u.v[0] = r0 | (r1 << 31);
        u.v[1] = (r1 >> 1) | (r2 << 30);
        u.v[2] = (r2 >> 2) | (r3 << 29);
        u.v[3] = r3 >> 3;
        return u;
    }
    uint32_t d[8];
    t = (uint64_t)a.d[0] * w.d[0];
    d[0] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[1], t >> 31);
    t = mad32(a.d[1], w.d[0], t);
    d[1] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[2], t >> 31);
    t = mad32(a.d[1], w.d[1], t);
    t = mad32(a.d[2], w.d[0], t);
    d[2] = (uint32_t)t & kM31;
    t = mad32(a.d[0], w.d[3], t >> 31);
    t = mad32(a.d[1], w.d[2], t);
    t = mad32(a.d[2], w.d[1], t);
    t = mad32(a.d[3], w.d[0], t);
    d[3] = (uint32_t)t & kM31;
    t = mad32(a.d[1], w.d[3], t >> 31);
    t = mad32(a.d[2], w.d[2], t);
    t = mad32(a.d[3], w.d[1], t);
    d[4] = (uint32_t)t & kM31;
    t = mad32(a.d[2], w.d[3], t >> 31);
    t = mad32(a.d[3], w.d[2], t);
    d[5] = (uint32_t)t & kM31;
    t = mad32(a.d[3], w.d[3], t >> 31);
    d[6] = (uint32_t)t & kM31;
    d[7] = (uint32_t)(t >> 31);
    // u = t + m q with m = -t q^-1 mod R, found digit by digit: m_i =
    // c_i (-q^-1 mod 2^31) mod 2^31 clears digit i of the running sum (the
    // unique m < R that clears the low four digits).  SPARSE: q1 = q2 = 0.
    const R31 &q = mq.q;
    const uint32_t qi0 = mq.qinv.d[0];
    uint32_t m[4], r[4];
    uint64_t c = d[0];
    m[0] = ((uint32_t)c * qi0) & kM31;
    c = mad32(m[0], q.d[0], c);
    c = (c >> 31) + d[1];
    if constexpr (!SPARSE) c = mad32(m[0], q.d[1], c);
    m[1] = ((uint32_t)c * qi0) & kM31;
    c = mad32(m[1], q.d[0], c);
    c = (c >> 31) + d[2];
    if constexpr (!SPARSE) {
        c = mad32(m[0], q.d[2], c);
        c = mad32(m[1], q.d[1], c);
    }
    m[2] = ((uint32_t)c * qi0) & kM31;
    c = mad32(m[2], q.d[0], c);
    c = mad32(m[0], q.d[3], (c >> 31) + d[3]);
    if constexpr (!SPARSE) {
        c = mad32(m[1], q.d[2], c);
        c = mad32(m[2], q.d[1], c);
    }
    m[3] = ((uint32_t)c * qi0) & kM31;
    c = mad32(m[3], q.d[0], c);
    t = mad32(m[1], q.d[3], (c >> 31) + d[4]);
    if constexpr (!SPARSE) {
        t = mad32(m[2], q.d[2], t);
        t = mad32(m[3], q.d[1], t);
    }
    r[0] = (uint32_t)t & kM31;
    t = mad32(m[2], q.d[3], (t >> 31) + d[5]);
    if constexpr (!SPARSE) t = mad32(m[3], q.d[2], t);
    r[1] = (uint32_t)t & kM31;
    t = mad32(m[3], q.d[3], (t >> 31) + d[6]);
    r[2] = (uint32_t)t & kM31;
    r[3] = (uint32_t)(t >> 31) + d[7];  // u/R < 2q < 2^123: fits, no carry out
    // digits -> 32-bit limbs
    U128 u;
    u.v[0] = r[0] | (r[1] << 31);
    u.v[1] = (r[1] >> 1) | (r[2] << 30);
    u.v[2] = (r[2] >> 2) | (r[3] << 29);
    u.v[3] = r[3] >> 3;
    return u;
}

// x - y if x >= y else x (128-bit limbs).
BNTT_HD U128 sub_if_ge(const U128 &x, const U128 &y)
{
    uint32_t k, z0, z1, z2, z3;
    z0 = subb(x.v[0], y.v[0], 0, k);
    z1 = subb(x.v[1], y.v[1], k, k);
    z2 = subb(x.v[2], y.v[2], k, k);
    z3 = subb(x.v[3], y.v[3], k, k);
    U128 r;
    r.v[0] = k ? x.v[0] : z0;
    r.v[1] = k ? x.v[1] : z1;
    r.v[2] = k ? x.v[2] : z2;
    r.v[3] = k ? x.v[3] : z3;
    return r;
}

// Exact Montgomery product: canonical for a < 4q, w < q.
template <bool SPARSE = false>
BNTT_HD U128 montmul_r31(const U128 &a, const R31 &w, const MontQ &mq)
{
    return sub_if_ge(montmul_r31_lazy<SPARSE>(a, w, mq), mq.ql);
}

// Plain 128-bit add and subtract (no reduction).
BNTT_HD U128 add128l(const U128 &a, const U128 &b)
{
    uint32_t k;
    U128 r;
    r.v[0] = addc(a.v[0], b.v[0], 0, k);
    r.v[1] = addc(a.v[1], b.v[1], k, k);
    r.v[2] = addc(a.v[2], b.v[2], k, k);
    r.v[3] = addc(a.v[3], b.v[3], k, k);
    return r;
}
BNTT_HD U128 sub128l(const U128 &a, const U128 &b)
{
    uint32_t k;
    U128 r;
    r.v[0] = subb(a.v[0], b.v[0], 0, k);
    r.v[1] = subb(a.v[1], b.v[1], k, k);
    r.v[2] = subb(a.v[2], b.v[2], k, k);
    r.v[3] = subb(a.v[3], b.v[3], k, k);
    return r;
}

// Canonical (a + b) mod q and (a - b) mod q for a, b < q.
BNTT_HD U128 addmod_canon(const U128 &a, const U128 &b, const U128 &q)
{
    uint32_t k, s0, s1, s2, s3, z0, z1, z2, z3, bb;
    s0 = addc(a.v[0], b.v[0], 0, k);
    s1 = addc(a.v[1], b.v[1], k, k);
    s2 = addc(a.v[2], b.v[2], k, k);
    s3 = addc(a.v[3], b.v[3], k, k);
    z0 = subb(s0, q.v[0], 0, bb);
    z1 = subb(s1, q.v[1], bb, bb);
    z2 = subb(s2, q.v[2], bb, bb);
    z3 = subb(s3, q.v[3], bb, bb);
    U128 r;
    r.v[0] = bb ? s0 : z0;
    r.v[1] = bb ? s1 : z1;
    r.v[2] = bb ? s2 : z2;
    r.v[3] = bb ? s3 : z3;
    return r;
}

BNTT_HD U128 submod_canon(const U128 &a, const U128 &b, const U128 &q) { return submod_ref(a, b, q); }

// ---------------------------------------------------- literal reference Barrett

struct W128 {
    uint64_t lo, hi;
};

BNTT_HD W128 mul64w(uint64_t a, uint64_t b)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0, p01 = (uint64_t)a0 * b1;
    const uint64_t p10 = (uint64_t)a1 * b0, p11 = (uint64_t)a1 * b1;
    const uint64_t mid = (p00 >> 32) + (uint32_t)p01 + (uint32_t)p10;
    W128 r;
    r.lo = (mid << 32) | (uint32_t)p00;
    r.hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
    return r;
}

BNTT_HD W128 add128(W128 a, W128 b)
{
    W128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
    return r;
}

BNTT_HD W128 sub128(W128 a, W128 b)
{
    W128 r;
    r.lo = a.lo - b.lo;
    r.hi = a.hi - b.hi - (a.lo < b.lo ? 1u : 0u);
    return r;
}

BNTT_HD bool lt128(W128 a, W128 b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// ntt_scalar.c:106-128 (mul128): the cross sum wraps at 2^128.
BNTT_HD void mul128_ref(W128 &ch, W128 &cl, W128 a, W128 b)
{
    const W128 albl = mul64w(a.lo, b.lo), albh = mul64w(a.lo, b.hi);
    const W128 ahbl = mul64w(a.hi, b.lo), ahbh = mul64w(a.hi, b.hi);
    const W128 cx = add128(albh, ahbl);
    const W128 cx2 = add128(cx, W128{albl.hi, 0});
    ch = add128(ahbh, W128{cx2.hi, 0});
    cl = W128{albl.lo, cx2.lo};
}

// ntt_scalar.c:130-149 (mullo128).
BNTT_HD W128 mullo128_ref(W128 a, W128 b)
{
    const W128 albl = mul64w(a.lo, b.lo);
    const uint64_t cx = a.lo * b.hi + a.hi * b.lo;
    return W128{albl.lo, cx + albl.hi};
}

// ntt_scalar.c:151-171 (mulhi128).
BNTT_HD W128 mulhi128_ref(W128 a, W128 b)
{
    const uint64_t alblh = mul64w(a.lo, b.lo).hi;
    const W128 albh = mul64w(a.lo, b.hi), ahbl = mul64w(a.hi, b.lo), ahbh = mul64w(a.hi, b.hi);
    const W128 cx = add128(albh, ahbl);
    const W128 cx2 = add128(cx, W128{alblh, 0});
    return add128(ahbh, W128{cx2.hi, 0});
}

// ntt_scalar.c:173-199 (mulmod128), MBITS = 124: shifts 122 and 6.
BNTT_HD U128 mulmod_ref(const U128 &a_, const U128 &b_, const U128 &m_, const U128 &mu_)
{
    const W128 a{lo64(a_), hi64(a_)}, b{lo64(b_), hi64(b_)};
    const W128 m{lo64(m_), hi64(m_)}, mu{lo64(mu_), hi64(mu_)};
    W128 qh, ql;
    mul128_ref(qh, ql, a, b);
    const W128 tmpl0 = ql;
    // ql = (ql >> 122) | (qh << 6)
    W128 q;
    q.lo = (ql.hi >> 58) | (qh.lo << 6);
    q.hi = (qh.lo >> 58) | (qh.hi << 6);
    const uint64_t qhq = qh.hi >> 58;
    const W128 qmh1 = mulhi128_ref(q, mu);
    const uint64_t qmh2 = qhq * mu.lo;
    const W128 qmh = add128(qmh1, W128{qmh2, 0});
    const W128 qq{(qmh.lo >> 1) | (qmh.hi << 63), qmh.hi >> 1};
    const W128 qqml = mullo128_ref(qq, m);
    const W128 tmpl = sub128(tmpl0, qqml);
    const W128 r = lt128(tmpl, m) ? tmpl : sub128(tmpl, m);
    return make_u128(r.lo, r.hi);
}

// --------------------------------------------------------------------- dispatch

BNTT_HD bool below_2_124(const U128 &a, const U128 &b) { return ((a.v[3] | b.v[3]) >> 28) == 0; }

// Elementwise product: the radix-2^31 reduction when the modulus is of the
// fast form and both operands are below 2^124, the literal formula otherwise.
// The test is made per wave: a per-lane choice gets if-converted into code
// that computes both products for every lane (about 370 instead of 75 VALU
// ops), and keeping the literal path out of line as a call costs every wave a
// stack allocation, which throttles how many waves are in flight.
template <bool PM>
__device__ __forceinline__ U128 mulmod_elem(const U128 &a, const U128 &b, const Mod &md)
{
    const bool fast = PM && below_2_124(a, b);
    if (__all(fast)) return mulmod_r31(b, to_r31(a), md.c);
    return fast ? mulmod_r31(b, to_r31(a), md.c) : mulmod_ref(a, b, md.m, md.mu);
}

template <bool PM>
BNTT_HD U128 mulmod(const U128 &a, const U128 &b, const Mod &md)
{
    if (PM && below_2_124(a, b)) return mulmod_pm(a, b, md.c);
    return mulmod_ref(a, b, md.m, md.mu);
}

BNTT_HD U128 addmod(const U128 &a, const U128 &b, const Mod &md) { return addmod_ref(a, b, md.m, md.notm); }
BNTT_HD U128 submod(const U128 &a, const U128 &b, const Mod &md) { return submod_ref(a, b, md.m); }

}  // namespace bntt
