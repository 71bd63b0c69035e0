"""ctypes front end of oracle.c (the CPU restatement).  TEST INFRASTRUCTURE ONLY.

Arrays of 128-bit residues are numpy ``uint64`` arrays of shape ``(n, 2)`` in
the native little-endian layout of ``unsigned __int128``: column 0 is the low
word, column 1 the high word.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "oracle.c")
_LIB = os.path.join(_HERE, "liboracle.so")


def build(force: bool = True) -> str:
    """Compile oracle.c into oracle/liboracle.so with gcc."""
    if force or not os.path.exists(_LIB):
        subprocess.run(
            ["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-fPIC", "-shared", _SRC, "-o", _LIB],
            check=True,
        )
    return _LIB


_P = ctypes.c_void_p


def _ptr(a: np.ndarray) -> int:
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class _Keep:
    """Holds the temporary scalar arrays of one call alive until it returns."""

    def __init__(self):
        self.arrays = []

    def __call__(self, v: int) -> int:
        a = u128_scalar(v)
        self.arrays.append(a)
        return _ptr(a)


def u128_scalar(v: int) -> np.ndarray:
    return np.array([[v & (2**64 - 1), (v >> 64) & (2**64 - 1)]], dtype=np.uint64)


def to_ints(a: np.ndarray) -> list[int]:
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in a]


def from_ints(vals) -> np.ndarray:
    vals = list(vals)
    out = np.empty((len(vals), 2), dtype=np.uint64)
    for i, v in enumerate(vals):
        out[i, 0] = v & (2**64 - 1)
        out[i, 1] = (v >> 64) & (2**64 - 1)
    return out


class Oracle:
    """Thin wrapper: every method takes and returns (n, 2) uint64 arrays."""

    def __init__(self, path: str | None = None):
        path = path or _LIB
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        L = self.lib
        for name in ("or_addmod128_p", "or_submod128_p"):
            getattr(L, name).argtypes = [_P] * 4
        L.or_mulmod128_p.argtypes = [_P] * 5
        L.or_ntt.argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, _P, _P]
        L.or_ntt_batch.argtypes = [_P, _P, ctypes.c_int, ctypes.c_long, _P, _P, _P, _P, _P]
        L.or_vvadd.argtypes = [_P, _P, _P, ctypes.c_size_t, _P]
        L.or_vvsub.argtypes = [_P, _P, _P, ctypes.c_size_t, _P]
        L.or_vvmul.argtypes = [_P, _P, _P, ctypes.c_size_t, _P, _P]
        L.or_axpy.argtypes = [_P, _P, _P, _P, ctypes.c_size_t, _P, _P]
        L.or_ntt_exact.argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, _P]
        L.or_intt_exact.argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, _P, _P]
        L.or_polymul_schoolbook.argtypes = [_P, _P, _P, ctypes.c_size_t, _P]
        L.or_splitmix_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.or_splitmix_at.restype = ctypes.c_uint64
        L.or_uniform.argtypes = [_P, ctypes.c_size_t, _P, ctypes.c_uint64]
        L.or_polymul_schoolbook_range.argtypes = [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t,
                                                  ctypes.c_size_t]

    # scalar primitives ------------------------------------------------------
    def addmod(self, a: int, b: int, m: int) -> int:
        k = _Keep()
        r = np.zeros((1, 2), np.uint64)
        self.lib.or_addmod128_p(_ptr(r), k(a), k(b), k(m))
        return to_ints(r)[0]

    def submod(self, a: int, b: int, m: int) -> int:
        k = _Keep()
        r = np.zeros((1, 2), np.uint64)
        self.lib.or_submod128_p(_ptr(r), k(a), k(b), k(m))
        return to_ints(r)[0]

    def mulmod(self, a: int, b: int, m: int, mu: int) -> int:
        k = _Keep()
        r = np.zeros((1, 2), np.uint64)
        self.lib.or_mulmod128_p(
            _ptr(r), k(a), k(b), k(m), k(mu)
        )
        return to_ints(r)[0]

    # kernels -------------------------------------------------------------------
    def ntt(self, x: np.ndarray, tw: np.ndarray, m: int, mu: int) -> np.ndarray:
        """Forward network of ntt_scalar.c:582-627 on one transform or a batch
        (x of shape (count*N, 2) with tw of shape (N, 2))."""
        k = _Keep()
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
        tw = np.ascontiguousarray(tw, dtype=np.uint64).reshape(-1, 2)
        n = tw.shape[0]
        log2n = n.bit_length() - 1
        assert n == 1 << log2n and n >= 2 and x.shape[0] % n == 0
        y = np.empty_like(x)
        as0 = np.empty((n, 2), np.uint64)
        as1 = np.empty((n, 2), np.uint64)
        self.lib.or_ntt_batch(
            _ptr(y), _ptr(x), log2n, x.shape[0] // n, k(m), k(mu),
            _ptr(tw), _ptr(as0), _ptr(as1),
        )
        return y

    def vvadd(self, a, b, m):
        k = _Keep()
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        c = np.empty_like(a)
        self.lib.or_vvadd(_ptr(c), _ptr(a), _ptr(b), a.shape[0], k(m))
        return c

    def vvsub(self, a, b, m):
        k = _Keep()
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        c = np.empty_like(a)
        self.lib.or_vvsub(_ptr(c), _ptr(a), _ptr(b), a.shape[0], k(m))
        return c

    def vvmul(self, a, b, m, mu):
        k = _Keep()
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        c = np.empty_like(a)
        self.lib.or_vvmul(_ptr(c), _ptr(a), _ptr(b), a.shape[0], k(m), k(mu))
        return c

    def axpy(self, a, b, x: int, m, mu):
        k = _Keep()
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        c = np.empty_like(a)
        self.lib.or_axpy(
            _ptr(c), _ptr(a), _ptr(b), k(x), a.shape[0], k(m),
            k(mu),
        )
        return c

    # exact extensions (polynomial path) --------------------------------------------
    def ntt_exact(self, x, tw, q: int):
        k = _Keep()
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
        tw = np.ascontiguousarray(tw, dtype=np.uint64).reshape(-1, 2)
        n = x.shape[0]
        log2n = n.bit_length() - 1
        y = np.empty_like(x)
        as0, as1 = np.empty_like(x), np.empty_like(x)
        self.lib.or_ntt_exact(_ptr(y), _ptr(x), log2n, k(q), _ptr(tw), _ptr(as0), _ptr(as1))
        return y

    def intt_exact(self, y, itw, q: int):
        k = _Keep()
        y = np.ascontiguousarray(y, dtype=np.uint64).reshape(-1, 2)
        itw = np.ascontiguousarray(itw, dtype=np.uint64).reshape(-1, 2)
        n = y.shape[0]
        log2n = n.bit_length() - 1
        x = np.empty_like(y)
        as0, as1 = np.empty_like(y), np.empty_like(y)
        inv2 = (q + 1) // 2
        self.lib.or_intt_exact(
            _ptr(x), _ptr(y), log2n, k(q), _ptr(itw), k(inv2), _ptr(as0), _ptr(as1)
        )
        return x

    def polymul_schoolbook(self, a, b, q: int, threads: int = 1):
        """Schoolbook negacyclic product; with threads > 1 the output
        coefficients are split into ranges computed concurrently."""
        k = _Keep()
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        c = np.empty_like(a)
        n = a.shape[0]
        if threads <= 1 or n < 256:
            self.lib.or_polymul_schoolbook(_ptr(c), _ptr(a), _ptr(b), n, k(q))
            return c
        from concurrent.futures import ThreadPoolExecutor

        qk = k(q)
        step = -(-n // (4 * threads))

        def part(k0):
            self.lib.or_polymul_schoolbook_range(_ptr(c), _ptr(a), _ptr(b), n, qk, k0, min(n, k0 + step))

        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(part, range(0, n, step)))
        return c


    def splitmix_at(self, seed: int, j: int) -> int:
        return self.lib.or_splitmix_at(seed, j)

    def uniform(self, n: int, q: int, seed: int):
        k = _Keep()
        out = np.empty((n, 2), dtype=np.uint64)
        self.lib.or_uniform(_ptr(out), n, k(q), seed)
        return out

    def polymul_ntt(self, a, b, q: int):
        """Negacyclic product through the exact networks: or_ntt_exact of a and
        b with the table tw[2^s + k] = psi_s^(2 brv_s(k) + 1) (psi_s of order
        2^(s+2)), pointwise product, or_intt_exact with the inverse table.  For
        a prime q with 2N | q - 1; agrees with polymul_schoolbook."""
        a, b = (np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 2) for v in (a, b))
        n = a.shape[0]
        L = n.bit_length() - 1
        tw, itw = negacyclic_tables(q, L)
        fa, fb = self.ntt_exact(a, from_ints(tw), q), self.ntt_exact(b, from_ints(tw), q)
        prod = from_ints([x * y % q for x, y in zip(to_ints(fa), to_ints(fb))])
        return self.intt_exact(prod, from_ints(itw), q)


def two_adic_root(q: int) -> tuple[int, int]:
    """(w, v): 2^v is the largest power of two dividing q - 1 and w is a root of
    unity of order exactly 2^v (g^((q-1)/2^v) for the least quadratic
    non-residue g; q prime)."""
    v = ((q - 1) & -(q - 1)).bit_length() - 1
    g = 2
    while pow(g, (q - 1) // 2, q) != q - 1:
        g += 1
    return pow(g, (q - 1) >> v, q), v


def _brv(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def root_table(q: int, L: int, kind: str) -> list[int]:
    """Twiddle table of 2^L entries (entry 0 = 0): 'cyclic' w_s^brv_s(k) with
    w_s of order 2^(s+1), 'negacyclic' p_s^(2 brv_s(k)+1) with p_s of order
    2^(s+2), '*_inv' the entry-wise inverses; the stage roots are powers of the
    root of two_adic_root(q)."""
    w, v = two_adic_root(q)
    nega = kind.startswith("negacyclic")
    inv = kind.endswith("_inv")
    if L + (1 if nega else 0) > v:
        raise ValueError("q - 1 has too few factors of two")
    tw = [0] * (1 << L)
    for s in range(L):
        order = s + (2 if nega else 1)
        r = pow(w, 1 << (v - order), q)
        if inv:
            r = pow(r, (1 << order) - 1, q)
        for k in range(1 << s):
            e = 2 * _brv(k, s) + 1 if nega else _brv(k, s)
            tw[(1 << s) + k] = pow(r, e, q)
    return tw


def negacyclic_tables(q: int, L: int) -> tuple[list[int], list[int]]:
    """Forward table (negacyclic NTT in bit-reversed order through the forward
    network) and its entry-wise inverse, for N = 2^L."""
    w, v = two_adic_root(q)
    if L + 1 > v:
        raise ValueError("q - 1 has too few factors of two")
    n = 1 << L
    tw, itw = [0] * n, [0] * n
    for s in range(L):
        r = pow(w, 1 << (v - s - 2), q)
        ri = pow(r, q - 2, q)
        for k in range(1 << s):
            e = 2 * _brv(k, s) + 1
            tw[(1 << s) + k] = pow(r, e, q)
            itw[(1 << s) + k] = pow(ri, e, q)
    return tw, itw


_ORACLE: Oracle | None = None


def load() -> Oracle:
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
