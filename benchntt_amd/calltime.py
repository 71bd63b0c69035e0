"""The reference's per-call timing protocol, run in C (libbntt_calltime.so,
built from csrc/call_timing.c) over the host-array entry points: each call
timed with the wall clock, the mean of the last half reported, as
ntt_scalar.c:671-684 and blas_scalar.c:660-721 time the reference's kernels.
Timing the calls in C leaves the interpreter's per-call cost out, as a C
caller of the reference would."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import U128

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbntt_calltime.so")
BLAS_OPS = {"vvadd": 0, "vvsub": 1, "vvmul": 2, "axpy": 3}

_ct = None


def load() -> ctypes.CDLL:
    global _ct
    if _ct is None:
        _lib.load()  # libbntt.so first, so the timing library binds to the same copy
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run python benchntt_amd/build.py")
        lib = ctypes.CDLL(LIB_PATH)
        p, u = ctypes.c_void_p, ctypes.POINTER(U128)
        lib.bntt_calltime_ntt.argtypes = [p, p, p, ctypes.c_int, u, u, ctypes.c_int]
        lib.bntt_calltime_ntt.restype = ctypes.c_double
        lib.bntt_calltime_blas.argtypes = [ctypes.c_int, p, p, p, u, ctypes.c_size_t, u, u, ctypes.c_int]
        lib.bntt_calltime_blas.restype = ctypes.c_double
        _ct = lib
    return _ct


def _ptr(a: np.ndarray, out: bool = False) -> int:
    if not isinstance(a, np.ndarray) or not a.flags.c_contiguous or a.nbytes % 16:
        raise ValueError("residue arrays must be C-contiguous numpy arrays of 16-byte elements")
    if out and not a.flags.writeable:
        raise ValueError("output arrays must be writeable")
    return a.ctypes.data


def _result(ns: float) -> float:
    if ns < 0:
        _lib.check(int(-ns))
    return ns


def ntt_ns_per_call(y: np.ndarray, x: np.ndarray, tw: np.ndarray, modulus: int, mu: int, iters: int = 100) -> float:
    """Mean ns per bntt_ntt_fwd call over the last iters // 2 of `iters`
    (y = NTT(x), one transform of len(tw) points; y holds the last output)."""
    n = tw.nbytes // 16
    if n < 1 or n & (n - 1) or x.nbytes != tw.nbytes or y.nbytes != x.nbytes:
        raise ValueError("one transform: x, y and tw hold N residues, N a power of two")
    m, u = U128.of(modulus), U128.of(mu)
    return _result(load().bntt_calltime_ntt(_ptr(y, True), _ptr(x), _ptr(tw), n.bit_length() - 1,
                                            ctypes.byref(m), ctypes.byref(u), iters))


def blas_ns_per_call(op: str, c: np.ndarray, a: np.ndarray, b: np.ndarray, modulus: int, mu: int = 0,
                     x: int = 0, iters: int = 100) -> float:
    """Mean ns per call of one host-array BLAS op (vvadd, vvsub, vvmul,
    axpy) over the last iters // 2 of `iters`; c holds the last output."""
    n = a.nbytes // 16
    if b.nbytes != a.nbytes or c.nbytes != a.nbytes:
        raise ValueError("arrays differ in length")
    m, u, s = U128.of(modulus), U128.of(mu), U128.of(x)
    return _result(load().bntt_calltime_blas(BLAS_OPS[op], _ptr(c, True), _ptr(a), _ptr(b), ctypes.byref(s), n,
                                             ctypes.byref(m), ctypes.byref(u), iters))
