/*
 * benchntt_compat.h -- the reference's own call shapes over the GPU library.
 *
 * A C file written against the reference kernels keeps its calls unchanged:
 * it defines N (the reference does: `#define stages 10`, `#define N (1 <<
 * stages)`, ntt_scalar.c:21-22), includes this header instead of defining the
 * kernels, and links libbntt.so.  Mapping (files of the reference's src/):
 *   ntt(Y, X, modulus, mu, twiddles, as0, as1)  ntt_scalar.c:582  -> bntt_ntt_fwd
 *   vector_add(c, a, b, m)                      blas_scalar.c:578 -> bntt_vvadd
 *   vector_sub(c, a, b, m)                      blas_scalar.c:586 -> bntt_vvsub
 *   vector_mul(c, a, b, m, mu)                  blas_scalar.c:594 -> bntt_vvmul
 *   axpy(c, a, b, x, m, mu)                     blas_scalar.c:602 -> bntt_axpy
 *   addmod128 / submod128 / mulmod128           ntt_scalar.c:49-199 -> bntt_addmod128 ...
 *                                               (host; define BNTT_COMPAT_NO_PRIMITIVES to omit)
 * Sizes: the NTT size is the caller's N; the BLAS length is BNTT_BLAS_LEN,
 * 1024 by default as in the reference (blas_scalar.c:580).  Pointers are host
 * pointers and the calls are synchronous, as in the reference.  as0/as1 are
 * accepted and unused (the library owns its device scratch).
 *
 * Error behaviour: the reference kernels return void and cannot fail; these
 * wrappers keep the void signatures and abort with the library's message if
 * a call fails (no device, bad size), rather than return wrong data.
 *
 * Also provided, as in the reference's config.h:3-7 and the MODULUS/MU macros
 * of blas_scalar.c:23-30: uint128_t, LO64, HI64, INT128, MBITS, MODULUS, MU.
 */
#ifndef BENCHNTT_COMPAT_H
#define BENCHNTT_COMPAT_H

#include <stdio.h>
#include <stdlib.h>

#include "bntt.h"
#include "bntt_config.h"

typedef unsigned __int128 uint128_t;

#ifndef LO64
#define LO64(val) ((uint64_t)(val))
#endif
#ifndef HI64
#define HI64(val) ((uint64_t)((val) >> 64))
#endif
#ifndef INT128
#define INT128(hi, lo) (((uint128_t)(hi) << 64) | (uint128_t)(lo))
#endif
#ifndef MBITS
#define MBITS 124
#endif
#ifndef MODULUS
#define MODULUS INT128(BNTT_QB_HI, BNTT_QB_LO)
#endif
#ifndef MU
#define MU INT128(BNTT_MUB_HI, BNTT_MUB_LO)
#endif

#ifndef BNTT_BLAS_LEN
#define BNTT_BLAS_LEN 1024
#endif

static inline bntt_u128 bntt_compat_u128(uint128_t v)
{
    bntt_u128 r;
    r.lo = (uint64_t)v;
    r.hi = (uint64_t)(v >> 64);
    return r;
}

static inline void bntt_compat_check(int rc, const char *what)
{
    if (rc != BNTT_OK) {
        fprintf(stderr, "%s failed: %s\n", what, bntt_last_error());
        abort();
    }
}

static inline int bntt_compat_log2(unsigned long n)
{
    int l = 0;
    while ((1UL << l) < n) l++;
    return l;
}

#ifndef BNTT_COMPAT_NO_PRIMITIVES
/* The scalar primitives (ntt_scalar.c:49-199), bit-identical, on the host. */
static inline uint128_t bntt_compat_v(bntt_u128 v) { return ((uint128_t)v.hi << 64) | v.lo; }

static inline uint128_t addmod128(uint128_t a, uint128_t b, uint128_t m)
{
    return bntt_compat_v(bntt_addmod128(bntt_compat_u128(a), bntt_compat_u128(b), bntt_compat_u128(m)));
}

static inline uint128_t submod128(uint128_t a, uint128_t b, uint128_t m)
{
    return bntt_compat_v(bntt_submod128(bntt_compat_u128(a), bntt_compat_u128(b), bntt_compat_u128(m)));
}

static inline uint128_t mulmod128(uint128_t a, uint128_t b, uint128_t m, uint128_t mu)
{
    return bntt_compat_v(
        bntt_mulmod128(bntt_compat_u128(a), bntt_compat_u128(b), bntt_compat_u128(m), bntt_compat_u128(mu)));
}
#endif

#ifdef N
static inline void ntt(uint128_t *Y, uint128_t *X, uint128_t modulus, uint128_t mu, uint128_t *twiddles,
                       uint128_t *as0, uint128_t *as1)
{
    bntt_u128 m = bntt_compat_u128(modulus), u = bntt_compat_u128(mu);
    (void)as0;
    (void)as1;
    bntt_compat_check(bntt_ntt_fwd((bntt_u128 *)Y, (const bntt_u128 *)X, (const bntt_u128 *)twiddles,
                                   bntt_compat_log2((unsigned long)(N)), &m, &u),
                      "ntt");
}
#endif

static inline void vector_add(uint128_t *c, uint128_t *a, uint128_t *b, uint128_t m)
{
    bntt_u128 mm = bntt_compat_u128(m);
    bntt_compat_check(bntt_vvadd((bntt_u128 *)c, (const bntt_u128 *)a, (const bntt_u128 *)b, BNTT_BLAS_LEN, &mm),
                      "vector_add");
}

static inline void vector_sub(uint128_t *c, uint128_t *a, uint128_t *b, uint128_t m)
{
    bntt_u128 mm = bntt_compat_u128(m);
    bntt_compat_check(bntt_vvsub((bntt_u128 *)c, (const bntt_u128 *)a, (const bntt_u128 *)b, BNTT_BLAS_LEN, &mm),
                      "vector_sub");
}

static inline void vector_mul(uint128_t *c, uint128_t *a, uint128_t *b, uint128_t m, uint128_t mu)
{
    bntt_u128 mm = bntt_compat_u128(m), uu = bntt_compat_u128(mu);
    bntt_compat_check(
        bntt_vvmul((bntt_u128 *)c, (const bntt_u128 *)a, (const bntt_u128 *)b, BNTT_BLAS_LEN, &mm, &uu),
        "vector_mul");
}

static inline void axpy(uint128_t *c, uint128_t *a, uint128_t *b, uint128_t x, uint128_t m, uint128_t mu)
{
    bntt_u128 mm = bntt_compat_u128(m), uu = bntt_compat_u128(mu), xx = bntt_compat_u128(x);
    bntt_compat_check(
        bntt_axpy((bntt_u128 *)c, (const bntt_u128 *)a, (const bntt_u128 *)b, &xx, BNTT_BLAS_LEN, &mm, &uu),
        "axpy");
}

#endif /* BENCHNTT_COMPAT_H */
