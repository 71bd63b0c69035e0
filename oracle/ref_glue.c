/*
 * ref_glue.c -- pointer-only entry points around the reference's own kernels,
 * so that tests and the CPU-baseline leg of bench.py can call the compiled
 * reference (oracle/_ref/*.so, built by oracle/build_ref.sh) through ctypes.
 * TEST INFRASTRUCTURE ONLY: the product library never loads these.
 *
 * It is compiled into the same shared object as one reference source file and
 * selects the matching variant with -DREF_{NTT,BLAS}_{SCALAR,AVX2,AVX512}.
 * The declarations below repeat the reference's signatures:
 *   ntt()            ntt_scalar.c:582
 *   ntt_pre/ntt/ntt_post  ntt_avx2.c:507,566,651; ntt_avx512.c:386,461,524
 *   vector_add/sub/mul, axpy  blas_scalar.c:578-608, blas_avx2.c:448-523,
 *                             blas_avx512.c:401-476
 * and the AVX files' process-wide constants `one`, `negone` (and `t` for AVX2)
 * are set the way their main() sets them (ntt_avx512.c:531-532,
 * ntt_avx2.c:658-660, blas_avx2.c:526-529, blas_avx512.c:479-480).
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

#ifndef BNTT_REF_STAGES
#define BNTT_REF_STAGES 10
#endif
#define RN (1 << BNTT_REF_STAGES)

static double now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e9 + (double)ts.tv_nsec;
}

int bntt_ref_log2n(void) { return BNTT_REF_STAGES; }

/* The reference's timing protocols take the mean of the last half of their
 * calls (ntt_scalar.c:670-683: 100 calls, mean of the last 50; the BLAS
 * drivers likewise, blas_scalar.c:660-721, with 1000 and the last 500 for the
 * AVX-512 add and sub, blas_avx512.c:560-590). */
static double mean_last_half(double *t, int iters)
{
    double s = 0.0;
    for (int i = iters / 2; i < iters; i++) s += t[i];
    return s / (iters - iters / 2);
}

/* ------------------------------------------------------------------ NTT */
#if defined(REF_NTT_SCALAR)
void ntt(u128 *Y, u128 *X, u128 modulus, u128 mu, u128 *twiddles, u128 *as0, u128 *as1);

int bntt_ref_init(void) { return 0; }

/* Y, X, tw: RN native u128 each (lo word first).  Returns ns per call: the
 * mean of the last half of `iters` timed calls (after one warm call). */
double bntt_ref_ntt(u128 *Y, u128 *X, const u128 *m, const u128 *mu, u128 *tw, int iters)
{
    u128 *as0 = (u128 *)malloc(sizeof(u128) * RN), *as1 = (u128 *)malloc(sizeof(u128) * RN);
    double *t = (double *)malloc(sizeof(double) * (iters > 0 ? iters : 1));
    ntt(Y, X, *m, *mu, tw, as0, as1);
    for (int i = 0; i < iters; i++) {
        const double t0 = now_ns();
        ntt(Y, X, *m, *mu, tw, as0, as1);
        t[i] = now_ns() - t0;
    }
    const double dt = iters > 0 ? mean_last_half(t, iters) : 0.0;
    free(t);
    free(as0);
    free(as1);
    return dt;
}
#endif

#if defined(REF_NTT_AVX512) || defined(REF_NTT_AVX2)
#if defined(REF_NTT_AVX512)
typedef __m512i vec;
#define LANES 8
extern __m512i one, negone;
int bntt_ref_init(void)
{
    negone = _mm512_set1_epi64(-1);
    one = _mm512_set1_epi64(1);
    return 0;
}
void ntt_pre(uint64_t *X, uint64_t *twiddles, u128 modulus, u128 mu, vec *tempx, vec *twdh,
             vec *twdl, vec *mh, vec *ml, vec *muh, vec *mul, vec *permutehi, vec *permutelo);
void ntt(vec *tempx, vec *twdh, vec *twdl, vec mh, vec ml, vec muh, vec mul, vec permutehi,
         vec permutelo, vec *xh, vec *xl, vec *ash, vec *asl, vec *th, vec *tl, vec *as);
#else
typedef __m256i vec;
#define LANES 4
extern __m256i one, negone, t;
int bntt_ref_init(void)
{
    negone = _mm256_set1_epi64x(-1);
    one = _mm256_set1_epi64x(1);
    t = _mm256_set1_epi64x((long long)0x8000000000000000ULL);
    return 0;
}
void ntt_pre(uint64_t *X, uint64_t *twiddles, u128 modulus, u128 mu, vec *tempx, vec *twdh,
             vec *twdl, vec *mh, vec *ml, vec *muh, vec *mul);
void ntt(vec *tempx, vec *twdh, vec *twdl, vec mh, vec ml, vec muh, vec mul, vec *xh, vec *xl,
         vec *ash, vec *asl, vec *th, vec *tl, vec *as);
#endif
void ntt_post(uint64_t *Y, vec *as);

static vec *valloc_(size_t n)
{
    vec *p = (vec *)aligned_alloc(64, sizeof(vec) * (n ? n : 1));
    memset(p, 0, sizeof(vec) * (n ? n : 1));
    return p;
}

/* Y, X, tw in the reference's hi/lo-u64 layout (2*RN words each).  Runs
 * ntt_pre once per call (untimed, as in the reference driver), times ntt()
 * only, and returns ns per ntt(): the mean of the last half of `iters` timed
 * calls (after one warm call). */
double bntt_ref_ntt(uint64_t *Y, uint64_t *X, const u128 *m, const u128 *mu, uint64_t *tw, int iters)
{
    const size_t v = RN / LANES, ops = RN / (2 * LANES), nv = (2 * (size_t)RN) / LANES;
    vec *twdh = valloc_(BNTT_REF_STAGES * ops), *twdl = valloc_(BNTT_REF_STAGES * ops);
    vec *tempx = valloc_(nv), *as = valloc_(nv);
    vec *xh = valloc_(v), *xl = valloc_(v), *ash = valloc_(v), *asl = valloc_(v);
    vec *th = valloc_(v), *tl = valloc_(v);
    vec mh, ml, muh, mul;
    double dt = 0.0;
#if defined(REF_NTT_AVX512)
    vec ph, pl;
    ntt_pre(X, tw, *m, *mu, tempx, twdh, twdl, &mh, &ml, &muh, &mul, &ph, &pl);
    double *t = (double *)malloc(sizeof(double) * (iters > 0 ? iters : 1));
    ntt(tempx, twdh, twdl, mh, ml, muh, mul, ph, pl, xh, xl, ash, asl, th, tl, as);
    for (int i = 0; i < iters; i++) {
        const double t0 = now_ns();
        ntt(tempx, twdh, twdl, mh, ml, muh, mul, ph, pl, xh, xl, ash, asl, th, tl, as);
        t[i] = now_ns() - t0;
    }
#else
    double *t = (double *)malloc(sizeof(double) * (iters > 0 ? iters : 1));
    ntt_pre(X, tw, *m, *mu, tempx, twdh, twdl, &mh, &ml, &muh, &mul);
    ntt(tempx, twdh, twdl, mh, ml, muh, mul, xh, xl, ash, asl, th, tl, as);
    for (int i = 0; i < iters; i++) {
        const double t0 = now_ns();
        ntt(tempx, twdh, twdl, mh, ml, muh, mul, xh, xl, ash, asl, th, tl, as);
        t[i] = now_ns() - t0;
    }
#endif
    if (iters > 0) dt = mean_last_half(t, iters);
    free(t);
    ntt_post(Y, as);
    free(twdh); free(twdl); free(tempx); free(as);
    free(xh); free(xl); free(ash); free(asl); free(th); free(tl);
    return dt;
}
#endif

/* ----------------------------------------------------------------- BLAS */
#if defined(REF_BLAS_SCALAR)
void vector_add(u128 *c, u128 *a, u128 *b, u128 m);
void vector_sub(u128 *c, u128 *a, u128 *b, u128 m);
void vector_mul(u128 *c, u128 *a, u128 *b, u128 m, u128 mu);
void axpy(u128 *c, u128 *a, u128 *b, u128 x, u128 m, u128 mu);
u128 addmod128(u128 a, u128 b, u128 m);
u128 submod128(u128 a, u128 b, u128 m);
u128 mulmod128(u128 a, u128 b, u128 m, u128 mu);

int bntt_ref_init(void) { return 0; }

/* Scalar primitives by pointer (blas_scalar.c copies of ntt_scalar.c:49-199). */
void bntt_ref_addmod128(u128 *r, const u128 *a, const u128 *b, const u128 *m) { *r = addmod128(*a, *b, *m); }
void bntt_ref_submod128(u128 *r, const u128 *a, const u128 *b, const u128 *m) { *r = submod128(*a, *b, *m); }
void bntt_ref_mulmod128(u128 *r, const u128 *a, const u128 *b, const u128 *m, const u128 *mu)
{
    *r = mulmod128(*a, *b, *m, *mu);
}

/* op: 0 vvadd, 1 vvsub, 2 vvmul, 3 axpy.  Arrays: nchunks*1024 native u128.
 * Returns the elapsed ns of the whole loop. */
double bntt_ref_blas(int op, u128 *c, u128 *a, u128 *b, long nchunks, const u128 *m,
                     const u128 *mu, const u128 *x)
{
    double t0 = now_ns();
    for (long k = 0; k < nchunks; k++) {
        u128 *cc = c + k * 1024, *aa = a + k * 1024, *bb = b + k * 1024;
        switch (op) {
        case 0: vector_add(cc, aa, bb, *m); break;
        case 1: vector_sub(cc, aa, bb, *m); break;
        case 2: vector_mul(cc, aa, bb, *m, *mu); break;
        default: axpy(cc, aa, bb, *x, *m, *mu); break;
        }
    }
    return now_ns() - t0;
}

double bntt_ref_blas_repeat(int op, u128 *c, u128 *a, u128 *b, int iters, const u128 *m, const u128 *mu,
                            const u128 *x)
{
    double *t = (double *)malloc(sizeof(double) * (iters > 0 ? iters : 1));
    for (int i = 0; i < iters; i++) t[i] = bntt_ref_blas(op, c, a, b, 1, m, mu, x);
    double r = iters > 0 ? mean_last_half(t, iters) : 0.0;
    free(t);
    return r;
}
#endif

#if defined(REF_BLAS_AVX512) || defined(REF_BLAS_AVX2)
#if defined(REF_BLAS_AVX512)
typedef __m512i bvec;
extern __m512i one, negone;
#define SET1(v) _mm512_set1_epi64((long long)(v))
int bntt_ref_init(void)
{
    negone = _mm512_set1_epi64(-1);
    one = _mm512_set1_epi64(1);
    return 0;
}
#else
typedef __m256i bvec;
extern __m256i one, negone, t;
#define SET1(v) _mm256_set1_epi64x((long long)(v))
int bntt_ref_init(void)
{
    negone = _mm256_set1_epi64x(-1);
    one = _mm256_set1_epi64x(1);
    t = _mm256_set1_epi64x((long long)0x8000000000000000ULL);
    return 0;
}
#endif
void vector_add(bvec *c, bvec *a, bvec *b, bvec mh, bvec ml);
void vector_sub(bvec *c, bvec *a, bvec *b, bvec mh, bvec ml);
void vector_mul(bvec *c, bvec *a, bvec *b, bvec mh, bvec ml, bvec muh, bvec mul);
void axpy(bvec *c, bvec *a, bvec *b, bvec xh, bvec xl, bvec mh, bvec ml, bvec muh, bvec mul);

/* Arrays in hi/lo-u64 layout, 64-byte aligned, nchunks*2048 words each. */
double bntt_ref_blas(int op, uint64_t *c, uint64_t *a, uint64_t *b, long nchunks, const u128 *m,
                     const u128 *mu, const u128 *x)
{
    bvec mh = SET1(*m >> 64), ml = SET1(*m), muh = SET1(*mu >> 64), mul = SET1(*mu);
    bvec xh = SET1(*x >> 64), xl = SET1(*x);
    double t0 = now_ns();
    for (long k = 0; k < nchunks; k++) {
        bvec *cc = (bvec *)(c + k * 2048), *aa = (bvec *)(a + k * 2048), *bb = (bvec *)(b + k * 2048);
        switch (op) {
        case 0: vector_add(cc, aa, bb, mh, ml); break;
        case 1: vector_sub(cc, aa, bb, mh, ml); break;
        case 2: vector_mul(cc, aa, bb, mh, ml, muh, mul); break;
        default: axpy(cc, aa, bb, xh, xl, mh, ml, muh, mul); break;
        }
    }
    return now_ns() - t0;
}

double bntt_ref_blas_repeat(int op, uint64_t *c, uint64_t *a, uint64_t *b, int iters, const u128 *m,
                            const u128 *mu, const u128 *x)
{
    double *t = (double *)malloc(sizeof(double) * (iters > 0 ? iters : 1));
    for (int i = 0; i < iters; i++) t[i] = bntt_ref_blas(op, c, a, b, 1, m, mu, x);
    double r = iters > 0 ? mean_last_half(t, iters) : 0.0;
    free(t);
    return r;
}
#endif
