/*
 * oracle.c -- CPU restatement of the benchNTT hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * library in benchntt_amd/ never links, loads or calls it.
 *
 * It restates, in plain C with unsigned __int128, the scalar algorithm of the
 * reference (all citations are files of the reference's src/ directory):
 *   or_addmod128  ntt_scalar.c:49-79   (same text in blas_scalar.c)
 *   or_submod128  ntt_scalar.c:81-104
 *   or_mul128     ntt_scalar.c:106-128
 *   or_mullo128   ntt_scalar.c:130-149
 *   or_mulhi128   ntt_scalar.c:151-171
 *   or_mulmod128  ntt_scalar.c:173-199  (MBITS = 124, USE_MBIT124)
 *   or_ntt        ntt_scalar.c:582-627  (N a runtime argument instead of a macro)
 *   or_vv*, or_axpy  blas_scalar.c:578-608 (length a runtime argument)
 * The formulas are followed literally, including the reference's edge
 * behaviour: or_addmod128 returns a value off by 2^64 (or m itself) whenever
 * lo64(a+b) == lo64(m) and a+b >= m (the borrow of the high word is taken
 * from !c5, ntt_scalar.c:73).
 *
 * Extensions that the reference does not have (inverse NTT and negacyclic
 * polynomial multiplication over q_116) are at the end of the file; they use
 * an exact reduction written from the mathematics (x mod q by long division),
 * not the reference's Barrett, whose MBITS=116 shift schedule is wrong.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - all 7 reference NTT fixtures x_N/twd_N -> ver_N (N=16..1024) and the 4
 *     BLAS fixtures, bit-exact;
 *   - or_mulmod128 == (a*b) mod q by Python big integers on random and edge
 *     operands; or_addmod128/or_submod128 == (a+-b) mod q away from the edge
 *     above, and the edge itself against its closed form;
 *   - or_ntt against the compiled reference ntt() (oracle/_ref) at N up to 2^20
 *     when that build exists, and against an O(N^2) evaluation of the network's
 *     closed form (bit-reversed DFT) for root-of-unity twiddles over q_116;
 *   - inverse NTT and polymul against schoolbook negacyclic convolution.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef unsigned __int128 u128;

#define MBITS 124

#define LO64(v) ((uint64_t)(v))
#define HI64(v) ((uint64_t)((v) >> 64))
#define U128(hi, lo) ((((u128)(uint64_t)(hi)) << 64) | (u128)(uint64_t)(lo))

/* ntt_scalar.c:49-79 */
u128 or_addmod128(u128 a, u128 b, u128 m)
{
    uint64_t al = LO64(a), ah = HI64(a), bl = LO64(b), bh = HI64(b);
    uint64_t ml = LO64(m), mh = HI64(m);
    u128 s16 = (u128)al + (u128)bl;
    uint64_t t18 = (uint64_t)s16;
    unsigned c3 = (unsigned)(s16 >> 64) & 1u;
    u128 s18 = (u128)ah + (u128)bh + (u128)c3;
    uint64_t t19 = (uint64_t)s18;
    u128 cx = U128(t19, t18);
    u128 ax = (u128)ml - (u128)t18;
    uint64_t s17a = (uint64_t)ax;
    unsigned c5 = (unsigned)(ax >> 64) & 1u;
    uint64_t d4 = (uint64_t)0 - s17a;
    uint64_t d6 = t19 - mh - (uint64_t)(!c5);
    u128 dx = U128(d6, d4);
    u128 s18a = (u128)mh - (u128)t19 - (u128)c5;
    unsigned i14 = (unsigned)(s18a >> 64) & 1u;
    return i14 ? dx : cx;
}

/* ntt_scalar.c:81-104 */
u128 or_submod128(u128 a, u128 b, u128 m)
{
    uint64_t al = LO64(a), ah = HI64(a), bl = LO64(b), bh = HI64(b);
    u128 s16 = (u128)al - (u128)bl;
    uint64_t t30 = (uint64_t)s16;
    unsigned c1 = (unsigned)(s16 >> 64) & 1u;
    u128 s17 = (u128)ah - (u128)bh - (u128)c1;
    uint64_t t29 = (uint64_t)s17;
    unsigned i28 = (unsigned)(s17 >> 64) & 1u;
    u128 cx = U128(t29, t30);
    u128 cx2 = cx + m;
    return i28 ? cx2 : cx;
}

/* ntt_scalar.c:106-128: 256-bit product as (hi, lo). */
void or_mul128(u128 *ch, u128 *cl, u128 a, u128 b)
{
    uint64_t al = LO64(a), ah = HI64(a), bl = LO64(b), bh = HI64(b);
    u128 albl = (u128)al * bl, albh = (u128)al * bh;
    u128 ahbl = (u128)ah * bl, ahbh = (u128)ah * bh;
    u128 cx = albh + ahbl;
    u128 cx2 = cx + (albl >> 64);
    u128 chx = ahbh + (cx2 >> 64);
    *ch = chx;
    *cl = U128((uint64_t)cx2, (uint64_t)albl);
}

/* ntt_scalar.c:130-149: low 128 bits of a*b. */
u128 or_mullo128(u128 a, u128 b)
{
    uint64_t al = LO64(a), ah = HI64(a), bl = LO64(b), bh = HI64(b);
    u128 albl = (u128)al * bl;
    uint64_t cx = al * bh + ah * bl;
    uint64_t cx2 = cx + (uint64_t)(albl >> 64);
    return U128(cx2, (uint64_t)albl);
}

/* ntt_scalar.c:151-171: high 128 bits of a*b, with the reference's 128-bit
 * wrap of the cross sum albh + ahbl (exact for the Barrett operands used). */
u128 or_mulhi128(u128 a, u128 b)
{
    uint64_t al = LO64(a), ah = HI64(a), bl = LO64(b), bh = HI64(b);
    uint64_t alblh = (uint64_t)(((u128)al * bl) >> 64);
    u128 albh = (u128)al * bh, ahbl = (u128)ah * bl, ahbh = (u128)ah * bh;
    u128 cx = albh + ahbl;
    u128 cx2 = cx + alblh;
    return ahbh + (cx2 >> 64);
}

/* ntt_scalar.c:173-199: Barrett reduction with the MBITS = 124 shift schedule. */
u128 or_mulmod128(u128 a, u128 b, u128 m, u128 mu)
{
    u128 qh, ql;
    or_mul128(&qh, &ql, a, b);
    u128 tmpl = ql;
    u128 ql1 = ql >> (MBITS - 2);
    u128 ql2 = qh << (128 - (MBITS - 2));
    ql = ql2 | ql1;
    uint64_t qhq = (uint64_t)(qh >> (MBITS - 2));
    u128 qmh1 = or_mulhi128(ql, mu);
    uint64_t qmh2 = qhq * LO64(mu);
    u128 qmh = qmh1 + qmh2;
    u128 qq = qmh >> 1;
    u128 qqml = or_mullo128(qq, m);
    tmpl = tmpl - qqml;
    return (tmpl < m) ? tmpl : tmpl - m;
}

/* ntt_scalar.c:582-627: forward constant-geometry radix-2 network.
 * Stage s (t = 2^s), j in [0, N/2):
 *   p = mulmod(tw[t + (j mod t)], in[j + N/2]);
 *   out[2j] = addmod(in[j], p); out[2j+1] = submod(in[j], p).
 * as0/as1 are caller scratch of N elements each (as in the reference); X is
 * not modified.  log2n >= 1. */
void or_ntt(u128 *Y, const u128 *X, int log2n, const u128 *m_, const u128 *mu_,
            const u128 *tw, u128 *as0, u128 *as1)
{
    const u128 m = *m_, mu = *mu_;
    const size_t n = (size_t)1 << log2n, h = n / 2;
    const u128 *in = X;
    for (int s = 0; s < log2n; s++) {
        u128 *out = (s == log2n - 1) ? Y : ((s & 1) ? as1 : as0);
        const size_t t = (size_t)1 << s;
        for (size_t j = 0; j < h; j++) {
            u128 p = or_mulmod128(tw[t + (j & (t - 1))], in[j + h], m, mu);
            out[2 * j] = or_addmod128(in[j], p, m);
            out[2 * j + 1] = or_submod128(in[j], p, m);
        }
        in = out;
    }
}

/* Batched form of or_ntt: `count` transforms of size 2^log2n, contiguous. */
void or_ntt_batch(u128 *Y, const u128 *X, int log2n, long count, const u128 *m,
                  const u128 *mu, const u128 *tw, u128 *as0, u128 *as1)
{
    const size_t n = (size_t)1 << log2n;
    for (long i = 0; i < count; i++)
        or_ntt(Y + (size_t)i * n, X + (size_t)i * n, log2n, m, mu, tw, as0, as1);
}

/* blas_scalar.c:578-608 with the length as an argument. */
void or_vvadd(u128 *c, const u128 *a, const u128 *b, size_t n, const u128 *m)
{
    for (size_t i = 0; i < n; i++) c[i] = or_addmod128(a[i], b[i], *m);
}

void or_vvsub(u128 *c, const u128 *a, const u128 *b, size_t n, const u128 *m)
{
    for (size_t i = 0; i < n; i++) c[i] = or_submod128(a[i], b[i], *m);
}

void or_vvmul(u128 *c, const u128 *a, const u128 *b, size_t n, const u128 *m,
              const u128 *mu)
{
    for (size_t i = 0; i < n; i++) c[i] = or_mulmod128(a[i], b[i], *m, *mu);
}

/* blas_scalar.c:602-608: c = a + x*b (the scalar multiplies b). */
void or_axpy(u128 *c, const u128 *a, const u128 *b, const u128 *x, size_t n,
             const u128 *m, const u128 *mu)
{
    for (size_t i = 0; i < n; i++)
        c[i] = or_addmod128(a[i], or_mulmod128(*x, b[i], *m, *mu), *m);
}

/* Scalar entry points by pointer, for callers that cannot pass __int128. */
void or_addmod128_p(u128 *r, const u128 *a, const u128 *b, const u128 *m)
{
    *r = or_addmod128(*a, *b, *m);
}
void or_submod128_p(u128 *r, const u128 *a, const u128 *b, const u128 *m)
{
    *r = or_submod128(*a, *b, *m);
}
void or_mulmod128_p(u128 *r, const u128 *a, const u128 *b, const u128 *m, const u128 *mu)
{
    *r = or_mulmod128(*a, *b, *m, *mu);
}

/* ------------------------------------------------------------------------- */
/* Extensions beyond the reference: exact arithmetic for the polynomial path.  */
/* ------------------------------------------------------------------------- */

/* (hi:lo) mod q for q < 2^127 by shift-subtract long division: exact, slow. */
static u128 mod256(u128 hi, u128 lo, u128 q)
{
    u128 r = hi % q;
    for (int i = 127; i >= 0; i--) {
        r = (r << 1) | ((lo >> i) & 1);
        if (r >= q) r -= q;
    }
    return r;
}

u128 or_mulmod_exact(u128 a, u128 b, u128 q)
{
    u128 hi, lo;
    or_mul128(&hi, &lo, a, b);
    return mod256(hi, lo, q);
}

static u128 addq(u128 a, u128 b, u128 q) { u128 s = a + b; return s >= q ? s - q : s; }
static u128 subq(u128 a, u128 b, u128 q) { return a >= b ? a - b : a + q - b; }

/* Forward network of or_ntt with exact canonical arithmetic mod q (q < 2^126).
 * With tw[2^s + k] = psi^(2*brv_s(k)+1) for psi a primitive 2^(s+2)-th ...
 * (table built by the caller) it computes the negacyclic NTT in bit-reversed
 * order. */
void or_ntt_exact(u128 *Y, const u128 *X, int log2n, const u128 *q_, const u128 *tw,
                  u128 *as0, u128 *as1)
{
    const u128 q = *q_;
    const size_t n = (size_t)1 << log2n, h = n / 2;
    const u128 *in = X;
    for (int s = 0; s < log2n; s++) {
        u128 *out = (s == log2n - 1) ? Y : ((s & 1) ? as1 : as0);
        const size_t t = (size_t)1 << s;
        for (size_t j = 0; j < h; j++) {
            u128 p = or_mulmod_exact(tw[t + (j & (t - 1))], in[j + h], q);
            out[2 * j] = addq(in[j], p, q);
            out[2 * j + 1] = subq(in[j], p, q);
        }
        in = out;
    }
}

/* Inverse of the forward network: stage s run backwards (s = L-1 .. 0):
 *   given out[2j], out[2j+1]:  in[j] = (o0 + o1)/2,  in[j+N/2] = (o0 - o1)/(2 tw).
 * itw[k] holds tw[k]^-1 and inv2 = 2^-1 mod q; the 1/2 factors are applied
 * per stage, so the result is X exactly. */
void or_intt_exact(u128 *X, const u128 *Y, int log2n, const u128 *q_, const u128 *itw,
                   const u128 *inv2_, u128 *as0, u128 *as1)
{
    const u128 q = *q_, inv2 = *inv2_;
    const size_t n = (size_t)1 << log2n, h = n / 2;
    const u128 *in = Y;
    for (int s = log2n - 1; s >= 0; s--) {
        u128 *out = (s == 0) ? X : ((s & 1) ? as1 : as0);
        const size_t t = (size_t)1 << s;
        for (size_t j = 0; j < h; j++) {
            u128 o0 = in[2 * j], o1 = in[2 * j + 1];
            u128 u = or_mulmod_exact(addq(o0, o1, q), inv2, q);
            u128 d = or_mulmod_exact(subq(o0, o1, q), inv2, q);
            out[j] = u;
            out[j + h] = or_mulmod_exact(d, itw[t + (j & (t - 1))], q);
        }
        in = out;
    }
}

/* Coefficients k0 <= k < k1 of the schoolbook negacyclic product
 * c = a*b mod (x^n + 1, q):  c[k] = sum_{i<=k} a_i b_{k-i} - sum_{i>k} a_i b_{n+k-i}.
 * Each output is independent, so callers may split [0, n) over threads. */
void or_polymul_schoolbook_range(u128 *c, const u128 *a, const u128 *b, size_t n, const u128 *q_,
                                 size_t k0, size_t k1)
{
    const u128 q = *q_;
    for (size_t k = k0; k < k1 && k < n; k++) {
        u128 acc = 0;
        for (size_t i = 0; i <= k; i++) acc = addq(acc, or_mulmod_exact(a[i], b[k - i], q), q);
        for (size_t i = k + 1; i < n; i++) acc = subq(acc, or_mulmod_exact(a[i], b[n + k - i], q), q);
        c[k] = acc;
    }
}

/* Schoolbook negacyclic product c = a*b mod (x^n + 1, q).  O(n^2). */
void or_polymul_schoolbook(u128 *c, const u128 *a, const u128 *b, size_t n, const u128 *q_)
{
    or_polymul_schoolbook_range(c, a, b, n, q_, 0, n);
}

/* ------------------------------------------------------------------------
 * Seeded uniform residues (this build's input generator; no counterpart in
 * the reference, whose inputs are fixed text files).  Word j of the
 * splitmix64 sequence from `seed` (Steele, Lea and Flood, OOPSLA 2014;
 * Vigna's public-domain splitmix64.c): z = seed + (j+1)*0x9E3779B97F4A7C15,
 * then two xor-shift-multiply rounds and a final xor-shift.
 */
uint64_t or_splitmix_at(uint64_t seed, uint64_t j)
{
    uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Residue i: the first draw t = 0, 1, ... below q of the words 2(t n + i)
 * (low) and 2(t n + i) + 1 (high), masked below 2^bitlen(q); the 64th draw,
 * if reached, is kept minus q. */
void or_uniform(u128 *out, size_t n, const u128 *q_, uint64_t seed)
{
    const u128 q = *q_;
    int bits = 0;
    while (bits < 128 && (q >> bits) != 0) bits++;
    const u128 mask = bits == 128 ? ~(u128)0 : (((u128)1 << bits) - 1);
    for (size_t i = 0; i < n; i++) {
        u128 v = 0;
        for (uint64_t t = 0; t < 64; t++) {
            const uint64_t j = 2 * (t * n + i);
            v = (((u128)or_splitmix_at(seed, j + 1) << 64) | or_splitmix_at(seed, j)) & mask;
            if (v < q) break;
            if (t == 63) v -= q;
        }
        out[i] = v;
    }
}
