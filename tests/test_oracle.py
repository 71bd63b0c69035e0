"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

1. The reference's golden vectors (ntt_scalar.c:685-701, blas_scalar.c:723-749).
2. Mathematics: big-integer closed forms for every primitive, the edge of the
   reference's addmod, and the network's closed form (bit-reversed DFT /
   negacyclic NTT) for root-of-unity twiddles, by O(N^2) evaluation.
3. The compiled reference itself (oracle/_ref) where that build exists.
"""
import random

import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints, to_ints, uniform

QN, MUN, QB, MUB = C.Q_NTT, C.MU_NTT, C.Q_BLAS, C.MU_BLAS


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024])
def test_ntt_golden(orc, fixtures, n):
    y = orc.ntt(fixtures[f"x_{n}"], fixtures[f"twd_{n}"], QN, MUN)
    assert np.array_equal(y, fixtures[f"ver_{n}"])


def test_blas_golden(orc, fixtures):
    a, b = fixtures["twd_1024"], fixtures["x_1024"]
    assert np.array_equal(orc.vvadd(a, b, QB), fixtures["vvadd_1024"])
    assert np.array_equal(orc.vvsub(a, b, QB), fixtures["vvsub_1024"])
    assert np.array_equal(orc.vvmul(a, b, QB, MUB), fixtures["vvmul_1024"])
    assert np.array_equal(orc.axpy(a, b, C.AXPY_SCALAR, QB, MUB), fixtures["axpy_1024"])


def test_blas_golden_closed_form(fixtures):
    """The BLAS fixtures are exact (a op b) mod q_B (checks the fixture data)."""
    a, b = to_ints(fixtures["twd_1024"]), to_ints(fixtures["x_1024"])
    s = C.AXPY_SCALAR
    assert to_ints(fixtures["vvmul_1024"]) == [x * y % QB for x, y in zip(a, b)]
    assert to_ints(fixtures["axpy_1024"]) == [(x + s * y) % QB for x, y in zip(a, b)]


def _edge_values(q):
    return [0, 1, 2, q - 1, q - 2, q // 2, (q + 1) // 2, 1 << 64, (1 << 64) - 1, (1 << 123), q - (1 << 64)]


@pytest.mark.parametrize("q,mu", [(QN, MUN), (QB, MUB)])
def test_mulmod_exact(orc, q, mu):
    """mulmod128 with mu = floor(2^251/q) is exactly a*b mod q for a, b < 2^124
    (the Barrett quotient is short by at most one)."""
    rng = random.Random(7)
    vals = _edge_values(q) + [rng.randrange(q) for _ in range(300)] + [rng.randrange(1 << 124) for _ in range(100)]
    vals += [(1 << 124) - 1, q, q + 1]
    for a in vals[:60]:
        for b in vals:
            assert orc.mulmod(a, b, q, mu) == a * b % q


@pytest.mark.parametrize("q", [QN, QB])
def test_addmod_submod(orc, q):
    rng = random.Random(11)
    vals = _edge_values(q) + [rng.randrange(q) for _ in range(200)]
    ml = q & (2**64 - 1)
    for a in vals[:80]:
        for b in vals:
            s = a + b
            if (s & (2**64 - 1)) == ml and s >= q:
                continue  # the reference's edge, pinned below
            assert orc.addmod(a, b, q) == s % q
            assert orc.submod(a, b, q) == (a - b) % q


@pytest.mark.parametrize("q", [QN, QB])
def test_addmod_reference_edge(orc, q):
    """ntt_scalar.c:73 takes the high-word borrow from !c5, so when
    lo64(a+b) == lo64(q) and a+b >= q the result is a+b-q-2^64 (and q itself
    when a+b == q).  Closed form checked on constructed operands."""
    rng = random.Random(3)
    ml = q & (2**64 - 1)
    cases = [(q - 1, 1), (1, q - 1), (q // 2, q - q // 2)]
    for _ in range(50):
        hi = rng.randrange(q >> 64, 2 * (q >> 64))  # a+b has lo word = ml and hi >= hi(q)
        s = (hi << 64) | ml
        a = rng.randrange(max(0, s - q + 1), min(q, s + 1))
        cases.append((a, s - a))
    for a, b in cases:
        s = a + b
        assert (s & (2**64 - 1)) == ml and s >= q
        expect = q if s == q else s - q - (1 << 64)
        assert orc.addmod(a, b, q) == expect


def _root_tables(q, L, negacyclic):
    g = 3
    while pow(g, (q - 1) // 2, q) == 1:
        g += 1
    w17 = pow(g, (q - 1) >> 17, q)

    def brv(x, b):
        return int(format(x, f"0{b}b")[::-1], 2) if b else 0

    n = 1 << L
    tw = [0] * n
    for s in range(L):
        if negacyclic:
            r = pow(w17, (1 << 17) >> (s + 2), q)
            for k in range(1 << s):
                tw[(1 << s) + k] = pow(r, 2 * brv(k, s) + 1, q)
        else:
            r = pow(w17, (1 << 17) >> (s + 1), q)
            for k in range(1 << s):
                tw[(1 << s) + k] = pow(r, brv(k, s), q)
    return tw, w17, brv


@pytest.mark.parametrize("L", [1, 2, 3, 5, 7])
def test_network_closed_form(orc, L):
    """With root-of-unity tables the network is the DFT in bit-reversed order
    (cyclic) or the negacyclic NTT: pins data flow, twiddle index and output
    order against O(N^2) evaluation, over q_116 (the only modulus with roots of
    order 2^17)."""
    q = C.Q_116
    n = 1 << L
    rng = random.Random(L)
    x = [rng.randrange(q) for _ in range(n)]
    for nega in (False, True):
        tw, w17, brv = _root_tables(q, L, nega)
        y = to_ints(orc.ntt_exact(from_ints(x), from_ints(tw), q))
        if nega:
            psi = pow(w17, (1 << 17) // (2 * n), q)
            ref = [sum(x[j] * pow(psi, j * (2 * k + 1), q) for j in range(n)) % q for k in range(n)]
        else:
            w = pow(w17, (1 << 17) // n, q)
            ref = [sum(x[j] * pow(w, j * k, q) for j in range(n)) % q for k in range(n)]
        assert y == [ref[brv(i, L)] for i in range(n)]


def test_reference_network_equals_exact_network(orc):
    """or_ntt (reference formulas, q_N) equals the exact-arithmetic network on
    random data: nothing in the Barrett path changes the values."""
    for L in (4, 9, 12):
        x = uniform(1 << L, QN, 100 + L)
        tw = uniform(1 << L, QN, 200 + L)
        assert np.array_equal(orc.ntt(x, tw, QN, MUN), orc.ntt_exact(x, tw, QN))


@pytest.mark.parametrize("n", [8, 64, 256])
def test_inverse_and_polymul(orc, n):
    """Inverse network round trip and negacyclic convolution theorem against
    schoolbook multiplication mod (x^n + 1, q_116)."""
    q = C.Q_116
    L = n.bit_length() - 1
    tw, _, _ = _root_tables(q, L, True)
    itw = [pow(t, q - 2, q) if t else 0 for t in tw]
    a, b = uniform(n, q, 1), uniform(n, q, 2)
    fa, fb = orc.ntt_exact(a, from_ints(tw), q), orc.ntt_exact(b, from_ints(tw), q)
    assert np.array_equal(orc.intt_exact(fa, from_ints(itw), q), a)
    prod = from_ints([x * y % q for x, y in zip(to_ints(fa), to_ints(fb))])
    c = orc.intt_exact(prod, from_ints(itw), q)
    assert np.array_equal(c, orc.polymul_schoolbook(a, b, q))
    # schoolbook itself against Python big integers
    ai, bi = to_ints(a), to_ints(b)
    ref = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                ref[k] += ai[i] * bi[j]
            else:
                ref[k - n] -= ai[i] * bi[j]
    assert to_ints(c) == [v % q for v in ref]


def test_uniform_residues_deterministic():
    a, b = uniform(1000, QN, 5), uniform(1000, QN, 5)
    assert np.array_equal(a, b)
    assert max(to_ints(a)) < QN
    assert not np.array_equal(a, uniform(1000, QN, 6))


# ------------------------------------------------------ the compiled reference

ref = pytest.importorskip("oracle.ref")


def _need(name):
    if not ref.available(name):
        pytest.skip(f"oracle/_ref/{name} not built (no reference sources here)")


@pytest.mark.parametrize("L", [4, 8, 10, 13, 16, 18])
def test_oracle_matches_reference_ntt(orc, L):
    _need(f"ref_ntt_scalar_L{L}.so")
    x, tw = uniform(1 << L, QN, L), uniform(1 << L, QN, 50 + L)
    y, _ = ref.ntt(x, tw, QN, MUN, "scalar")
    assert np.array_equal(y, orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("isa", ["avx2", "avx512"])
@pytest.mark.parametrize("L", [4, 5, 10, 11])
def test_oracle_matches_reference_avx_ntt(orc, isa, L):
    """The reference's AVX2 / AVX-512 ntt() at even and odd sizes (the
    benchmark uses them at every size of its sweeps)."""
    if isa == "avx512" and ref.best_isa() != "avx512":
        pytest.skip("host has no AVX-512")
    _need(f"ref_ntt_{isa}_L{L}.so")
    x, tw = uniform(1 << L, QN, L), uniform(1 << L, QN, 20 + L)
    y, _ = ref.ntt(x, tw, QN, MUN, isa)
    assert np.array_equal(y, orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("isa", ["scalar", "avx2", "avx512"])
def test_reference_amd_build_matches(orc, isa):
    """The reference objects built with its AMD configuration (clang
    -march=znver4) give the same outputs as the oracle, at an odd and an even
    size, and the BLAS protocol entry runs."""
    if ref.flavour() != "amd":
        pytest.skip("no znver4 objects, or the host lacks znver4's ISA")
    for L in (7, 12):
        _need(f"ref_ntt_{isa}_amd_L{L}.so")
        x, tw = uniform(1 << L, QN, 3 * L), uniform(1 << L, QN, 3 * L + 1)
        y, ns = ref.ntt(x, tw, QN, MUN, isa, iters=4, fl="amd")
        assert np.array_equal(y, orc.ntt(x, tw, QN, MUN)) and ns > 0
    a, b = uniform(1024, QB, 5), uniform(1024, QB, 6)
    ops = ref.BlasOperands(a, b, isa, fl="amd")
    ref.blas("vvmul", ops, QB, MUB)
    assert np.array_equal(ops.result(), orc.vvmul(a, b, QB, MUB))
    assert ref.blas_ns_per_element("axpy", ops, QB, MUB, C.AXPY_SCALAR, iters=10) > 0
    assert np.array_equal(ops.result(), orc.axpy(a, b, C.AXPY_SCALAR, QB, MUB))


def test_oracle_matches_reference_blas(orc):
    _need("ref_blas_scalar.so")
    a, b = uniform(4096, QB, 1), uniform(4096, QB, 2)
    ops = ref.BlasOperands(a, b, "scalar")
    for op, fn in (("vvadd", lambda: orc.vvadd(a, b, QB)), ("vvsub", lambda: orc.vvsub(a, b, QB)),
                   ("vvmul", lambda: orc.vvmul(a, b, QB, MUB)),
                   ("axpy", lambda: orc.axpy(a, b, C.AXPY_SCALAR, QB, MUB))):
        ref.blas(op, ops, QB, MUB, C.AXPY_SCALAR)
        assert np.array_equal(ops.result(), fn()), op


def test_oracle_matches_reference_primitives_on_edges(orc):
    """Scalar primitives, including the addmod edge and operands >= q, against
    the reference's own addmod128/submod128/mulmod128."""
    _need("ref_blas_scalar.so")
    rng = random.Random(9)
    q, mu = QB, MUB
    ml = q & (2**64 - 1)
    vals = _edge_values(q) + [rng.randrange(1 << 128) for _ in range(30)] + [q, q + 1, (1 << 124) - 1]
    pairs = [(a, b) for a in vals for b in vals]
    for _ in range(40):
        hi = rng.randrange(q >> 64, 2 * (q >> 64))
        s = (hi << 64) | ml
        a = rng.randrange(max(0, s - q + 1), min(q, s + 1))
        pairs.append((a, s - a))
    for a, b in pairs:
        assert orc.addmod(a, b, q) == ref.scalar_primitive("addmod128", a, b, q)
        assert orc.submod(a, b, q) == ref.scalar_primitive("submod128", a, b, q)
        assert orc.mulmod(a, b, q, mu) == ref.scalar_primitive("mulmod128", a, b, q, mu)


def _is_prime(n: int) -> bool:
    """Miller-Rabin with the first 16 prime bases (deterministic far beyond
    2^128 for composites of no special form; a check of the test moduli)."""
    if n < 2:
        return False
    bases = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53)
    for p in bases:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d, s = d // 2, s + 1
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


# moduli of the polynomial-multiply path: q_116, the largest prime below 2^122
# with 2^17 | q - 1, the 64-bit "Goldilocks" prime, and a tiny prime
POLY_MODULI = {"q116": (C.Q_116, 16), "q122": ((1 << 122) - 6160383, 16),
               "q64": ((1 << 64) - (1 << 32) + 1, 16), "q97": (97, 4)}


@pytest.mark.parametrize("name", sorted(POLY_MODULI))
def test_polymul_ntt_matches_schoolbook(orc, name):
    """The oracle's NTT-based negacyclic product (used for sizes where O(N^2)
    is too slow) equals schoolbook multiplication, for every modulus the GPU
    tests use; the moduli are prime with the stated 2-adicity."""
    q, lmax = POLY_MODULI[name]
    assert _is_prime(q)
    assert (q - 1) % (1 << (lmax + 1)) == 0
    rng = random.Random(q)
    for L in sorted({1, 2, 3, min(lmax, 6), min(lmax, 8)}):
        n = 1 << L
        a = from_ints([rng.randrange(q) for _ in range(n)])
        b = from_ints([rng.randrange(q) for _ in range(n)])
        assert np.array_equal(orc.polymul_ntt(a, b, q), orc.polymul_schoolbook(a, b, q))


def test_schoolbook_threaded_ranges(orc):
    """The threaded schoolbook (independent output ranges) equals the one-call
    schoolbook."""
    q = C.Q_116
    a, b = uniform(1024, q, 3), uniform(1024, q, 4)
    assert np.array_equal(orc.polymul_schoolbook(a, b, q, threads=8), orc.polymul_schoolbook(a, b, q))
