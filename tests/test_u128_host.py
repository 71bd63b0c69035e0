"""The device arithmetic of benchntt_amd/csrc/u128.hpp, compiled for the host,
against the oracle on random and edge operands (CPU only; catches arithmetic
mistakes before any GPU run)."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints, uniform

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def chk(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("u128") / "u128_host_check.so")
    src = os.path.join(HERE, "csrc", "u128_host_check.hip")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
                        src, "-o", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(out)
    lib.chk_op.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t] + [ctypes.c_void_p] * 2 + [
        ctypes.c_uint32]
    return lib


def run(lib, op, a, b, m, mu, c=0):
    out = np.zeros_like(a)
    mm, mmu = from_ints([m]), from_ints([mu])
    lib.chk_op(op, out.ctypes.data, a.ctypes.data, b.ctypes.data, a.shape[0], mm.ctypes.data, mmu.ctypes.data, c)
    return out


def operands(q, seed):
    rng = random.Random(seed)
    ml = q & (2**64 - 1)
    edges = [0, 1, q - 1, q - 2, q, q + 1, (1 << 124) - 1, 1 << 123, (1 << 64) - 1, 1 << 64, q // 2]
    big = [rng.randrange(1 << 128) for _ in range(20)]
    av, bv = [], []
    for x in edges + big:
        for y in edges + big:
            av.append(x)
            bv.append(y)
    for _ in range(200):  # the reference addmod's edge: lo64(a+b) == lo64(q), a+b >= q
        hi = rng.randrange(q >> 64, 2 * (q >> 64))
        s = (hi << 64) | ml
        x = rng.randrange(max(0, s - q + 1), min(q, s + 1))
        av.append(x)
        bv.append(s - x)
    a = np.concatenate([from_ints(av), uniform(20000, q, seed)])
    b = np.concatenate([from_ints(bv), uniform(20000, q, seed + 1)])
    return a, b


@pytest.mark.parametrize("q,mu", [(C.Q_NTT, C.MU_NTT), (C.Q_BLAS, C.MU_BLAS)])
def test_device_arith_matches_oracle(chk, orc, q, mu):
    a, b = operands(q, 17)
    c = (1 << 124) - q
    assert np.array_equal(run(chk, 0, a, b, q, mu), orc.vvadd(a, b, q))
    assert np.array_equal(run(chk, 1, a, b, q, mu), orc.vvsub(a, b, q))
    want = orc.vvmul(a, b, q, mu)
    assert np.array_equal(run(chk, 2, a, b, q, mu), want)
    assert np.array_equal(run(chk, 4, a, b, q, mu, c), want)  # fast path + fallback dispatch
    assert np.array_equal(run(chk, 5, a, b, q, mu), want)
    small = (a[:, 1] >> np.uint64(60) == 0) & (b[:, 1] >> np.uint64(60) == 0)
    assert np.array_equal(run(chk, 3, a[small], b[small], q, mu, c), want[small])
    assert np.array_equal(run(chk, 6, a[small], b[small], q, mu, c), want[small])
    # the fast-form add/sub (q's limbs 1..3 as constants): the reference's bits
    # for every operand, the addmod edge and operands >= 2^124 included
    assert np.array_equal(run(chk, 13, a, b, q, mu), orc.vvadd(a, b, q))
    assert np.array_equal(run(chk, 14, a, b, q, mu), orc.vvsub(a, b, q))


def test_device_arith_other_modulus(chk, orc):
    """A modulus that is not of the form 2^124 - c (and a mu that is not
    floor(2^251/m)) still follows the reference formulas literally."""
    q = C.Q_116
    mu = (1 << 235) // q
    a, b = operands(q, 5)
    assert np.array_equal(run(chk, 5, a, b, q, mu), orc.vvmul(a, b, q, mu))
    assert np.array_equal(run(chk, 0, a, b, q, mu), orc.vvadd(a, b, q))
    assert np.array_equal(run(chk, 1, a, b, q, mu), orc.vvsub(a, b, q))


def test_montgomery_and_canonical_ops(chk):
    """Exact Montgomery product a*w*2^-124 mod q and canonical add/sub, for the
    negacyclic polynomial path (modulus q_116), against big integers."""
    from benchntt_amd.residues import to_ints

    q = C.Q_116
    R = 1 << 124
    qinv = (-pow(q, -1, R)) % R
    rng = random.Random(21)
    edges = [0, 1, 2, q - 1, q - 2, q // 2, (1 << 64) - 1, 1 << 64]
    av = edges * len(edges) + [rng.randrange(q) for _ in range(20000)]
    bv = [e for e in edges for _ in edges] + [rng.randrange(q) for _ in range(20000)]
    a, b = from_ints(av), from_ints(bv)
    rinv = pow(R, -1, q)
    got = to_ints(run(chk, 7, a, b, q, qinv))
    assert got == [x * y * rinv % q for x, y in zip(av, bv)]
    assert to_ints(run(chk, 8, a, b, q, 0)) == [(x + y) % q for x, y in zip(av, bv)]
    assert to_ints(run(chk, 9, a, b, q, 0)) == [(x - y) % q for x, y in zip(av, bv)]


def test_lazy_montgomery(chk):
    """Lazy Montgomery product and butterfly (serving path): for inputs below
    4q and a canonical twiddle the product is below 2q and congruent to
    a*w*2^-124; the butterfly's outputs are below 4q and congruent to
    x + y*w*2^-124 and x - y*w*2^-124.  Also the exact product on inputs up to
    4q.  Moduli: q_116, the largest prime below 2^122 with 2^17 | q - 1, 97,
    and the largest odd q < 2^121 with zero middle digits; for those whose
    middle 31-bit digits are zero and q < 2^121 (the sparse kernel's range,
    where a column of product and reduction fits one 64-bit sum), the sparse
    reduction gives the same bits, also on operands whose digits are all at
    their maximum."""
    from benchntt_amd.residues import to_ints

    R = 1 << 124
    for q in (C.Q_116, (1 << 122) - 6160383, 97, (((1 << 28) - 1) << 93) + (1 << 31) - 1):
        qinv = (-pow(q, -1, R)) % R
        rinv = pow(R, -1, q)
        rng = random.Random(q)
        edges = [0, 1, q - 1, q, 2 * q - 1, 2 * q, 4 * q - 1, 3 * q + 5]
        # operands whose 31-bit digits are (nearly) all maximal below 4q
        edges += [v for v in ((1 << 124) - 1, (1 << 93) - 1, ((1 << 124) - 1) >> 2) if v < 4 * q]
        av = [e for e in edges for _ in edges] + [rng.randrange(4 * q) for _ in range(5000)]
        bv = edges * len(edges) + [rng.randrange(q) for _ in range(5000)]
        bv = [y % q for y in bv]  # the twiddle operand is canonical
        a, b = from_ints(av), from_ints(bv)
        lazy = to_ints(run(chk, 10, a, b, q, qinv))
        assert all(v < 2 * q and v % q == x * y * rinv % q for v, x, y in zip(lazy, av, bv))
        if (q >> 31) % (1 << 62) == 0 and q < (1 << 121):  # the sparse kernel's moduli
            assert to_ints(run(chk, 12, a, b, q, qinv)) == lazy
        assert to_ints(run(chk, 7, a, b, q, qinv)) == [x * y * rinv % q for x, y in zip(av, bv)]
        # butterfly with a one-digit twiddle w < min(q, 2^31)
        w = rng.randrange(min(q, 1 << 31))
        ys = from_ints([rng.randrange(4 * q) for _ in range(len(av))])
        out = np.zeros((2 * len(av), 2), dtype=np.uint64)
        mm, mmu = from_ints([q]), from_ints([qinv])
        chk.chk_op(11, out.ctypes.data, a.ctypes.data, ys.ctypes.data, len(av), mm.ctypes.data, mmu.ctypes.data, w)
        o = to_ints(out)
        for i, (x, y) in enumerate(zip(av, to_ints(ys))):
            p = y * w * rinv
            assert o[2 * i] < 4 * q and o[2 * i] % q == (x + p) % q
            assert o[2 * i + 1] < 4 * q and o[2 * i + 1] % q == (x - p) % q
