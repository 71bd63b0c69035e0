/*
 * bntt_config.h -- moduli, Barrett constants and 128-bit helper macros of the
 * benchNTT hot path (128-bit modular NTT and vvadd/vvsub/vvmul/axpy mod q).
 *
 * The constants are the ones the reference compiles in:
 *   NTT  : q_N = 2^124 - 719, mu_N = floor(2^251 / q_N) = 2^127 + 5752
 *          (ntt_scalar.c:632-633, ntt_avx2.c:694-695, ntt_avx512.c:568-569;
 *           the NTT drivers override the MODULUS/MU macros with these locals)
 *   BLAS : q_B = 2^124 - 59,  mu_B = floor(2^251 / q_B) = 2^127 + 472
 *          (MODULUS/MU macros, ntt_scalar.c:28-35, blas_scalar.c:23-30)
 *   axpy scalar s = 0x02E4B338A15CE95F_B1C09DAA097FEBD4 (blas_scalar.c:614)
 *   MBITS = 124 (ntt_scalar.c:17,29: USE_MBIT124)
 * plus the disabled alternative q_116 = 2^115 + 47*2^17 + 1 (ntt_scalar.c:36-43),
 * the one modulus of the three with 2^17-th roots of unity (q_116 - 1 = 2^17 * odd),
 * which the negacyclic polynomial-multiply path uses with its own exact reduction.
 *
 * Every constant is given as (hi, lo) 64-bit words; BNTT_U128(hi, lo) builds the
 * native unsigned __int128 value for C/C++ callers.
 */
#ifndef BNTT_CONFIG_H
#define BNTT_CONFIG_H

#include <stdint.h>

#ifndef BNTT_NO_U128_TYPEDEF
typedef unsigned __int128 bntt_uint128;
#endif

#define BNTT_LO64(val) ((uint64_t)(val))
#define BNTT_HI64(val) ((uint64_t)((val) >> 64))
#define BNTT_U128(hi, lo) \
    ((((unsigned __int128)(uint64_t)(hi)) << 64) | (unsigned __int128)(uint64_t)(lo))

#define BNTT_MBITS 124

/* NTT modulus q_N = 2^124 - 719 and mu_N = floor(2^251 / q_N). */
#define BNTT_QN_HI 0x0FFFFFFFFFFFFFFFULL
#define BNTT_QN_LO 0xFFFFFFFFFFFFFD31ULL
#define BNTT_MUN_HI 0x8000000000000000ULL
#define BNTT_MUN_LO 0x0000000000001678ULL

/* BLAS modulus q_B = 2^124 - 59 and mu_B = floor(2^251 / q_B). */
#define BNTT_QB_HI 0x0FFFFFFFFFFFFFFFULL
#define BNTT_QB_LO 0xFFFFFFFFFFFFFFC5ULL
#define BNTT_MUB_HI 0x8000000000000000ULL
#define BNTT_MUB_LO 0x00000000000001D8ULL

/* axpy scalar of the reference BLAS driver. */
#define BNTT_AXPY_S_HI 0x02E4B338A15CE95FULL
#define BNTT_AXPY_S_LO 0xB1C09DAA097FEBD4ULL

/* q_116 = 2^115 + 47 * 2^17 + 1 (prime, q_116 - 1 = 2^17 * odd). */
#define BNTT_Q116_HI 0x0008000000000000ULL
#define BNTT_Q116_LO 0x00000000005E0001ULL

#endif /* BNTT_CONFIG_H */
