"""Moduli and constants of the hot path (Python ints).  Same values as
include/bntt_config.h; see that header for their sources in the reference."""

MBITS = 124

# NTT: q_N = 2^124 - 719, mu_N = floor(2^251 / q_N)   (ntt_scalar.c:632-633)
Q_NTT = (1 << 124) - 719
MU_NTT = (1 << 251) // Q_NTT

# BLAS: q_B = 2^124 - 59, mu_B = floor(2^251 / q_B)   (blas_scalar.c:23-30)
Q_BLAS = (1 << 124) - 59
MU_BLAS = (1 << 251) // Q_BLAS

# axpy scalar of the reference BLAS driver (blas_scalar.c:614)
AXPY_SCALAR = (0x02E4B338A15CE95F << 64) | 0xB1C09DAA097FEBD4

# q_116 = 2^115 + 47 * 2^17 + 1: prime, q_116 - 1 = 2^17 * odd  (ntt_scalar.c:36-43)
Q_116 = (1 << 115) + 47 * (1 << 17) + 1

assert MU_NTT == (1 << 127) + 5752 and MU_BLAS == (1 << 127) + 472
