/*
 * call_timing.c -- the reference's own timing protocol, run in C against this
 * library's host-array entry points (the calls a C program written against
 * the reference makes after it links libbntt.so in place of the reference's
 * kernels).
 *
 * The reference times each call of its kernels with the wall clock and
 * reports the mean of the last half: ntt() in ntt_scalar.c:671-684 (100
 * calls, mean of the last 50, divided by N log2(N) / 2 butterflies), the BLAS
 * calls in blas_scalar.c:660-721 (100 calls, mean of the last 50, divided by
 * the 1024 elements).  Timing the calls here, rather than from an interpreter,
 * leaves out nothing but the caller's own loop, as in the reference.
 *
 * Built by benchntt_amd/build.py into benchntt_amd/libbntt_calltime.so, linked
 * against libbntt.so.  It is a measurement harness: the library itself does
 * not use it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

#include "bntt.h"

static double now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e9 + (double)ts.tv_nsec;
}

static double mean_last_half(const double *t, int iters)
{
    double s = 0.0;
    for (int i = iters / 2; i < iters; i++) s += t[i];
    return s / (iters - iters / 2);
}

/* `iters` calls of bntt_ntt_fwd(Y, X, tw, log2n, m, mu), each timed; returns
 * the mean ns per call of the last iters / 2, or -(error code) if a call
 * fails (bntt_last_error() then says why).  Y holds the last call's output. */
double bntt_calltime_ntt(bntt_u128 *Y, const bntt_u128 *X, const bntt_u128 *tw, int log2n, const bntt_u128 *m,
                         const bntt_u128 *mu, int iters)
{
    if (iters < 2) return -(double)BNTT_EINVAL;
    double *t = (double *)malloc(sizeof(double) * (size_t)iters);
    if (!t) return -(double)BNTT_EINVAL;
    for (int i = 0; i < iters; i++) {
        const double t0 = now_ns();
        const int rc = bntt_ntt_fwd(Y, X, tw, log2n, m, mu);
        t[i] = now_ns() - t0;
        if (rc != BNTT_OK) {
            free(t);
            return -(double)rc;
        }
    }
    const double r = mean_last_half(t, iters);
    free(t);
    return r;
}

/* The same for one BLAS op over n elements: op 0 vector_add (bntt_vvadd),
 * 1 vector_sub, 2 vector_mul, 3 axpy (x: its scalar); mu is read by 2 and 3. */
double bntt_calltime_blas(int op, bntt_u128 *c, const bntt_u128 *a, const bntt_u128 *b, const bntt_u128 *x,
                          size_t n, const bntt_u128 *m, const bntt_u128 *mu, int iters)
{
    if (iters < 2 || op < 0 || op > 3) return -(double)BNTT_EINVAL;
    double *t = (double *)malloc(sizeof(double) * (size_t)iters);
    if (!t) return -(double)BNTT_EINVAL;
    for (int i = 0; i < iters; i++) {
        int rc;
        const double t0 = now_ns();
        switch (op) {
        case 0: rc = bntt_vvadd(c, a, b, n, m); break;
        case 1: rc = bntt_vvsub(c, a, b, n, m); break;
        case 2: rc = bntt_vvmul(c, a, b, n, m, mu); break;
        default: rc = bntt_axpy(c, a, b, x, n, m, mu); break;
        }
        t[i] = now_ns() - t0;
        if (rc != BNTT_OK) {
            free(t);
            return -(double)rc;
        }
    }
    const double r = mean_last_half(t, iters);
    free(t);
    return r;
}
