"""Data formats and generated inputs of the path.

Text fixtures: the reference keeps its known-answer vectors as text, one
decimal 64-bit word per line, the high word then the low word of each residue
(data/x_N.txt, twd_N.txt, ver_N.txt for N = 16..1024, with no newline after
the last word, and data/<op>_1024.txt, with one; read by config.c:6-197 with
fscanf "%lu").  Sizes above
1024 repeat the 1024-point input and table (config.c:35-44, 101-134); there is
no expected output above 1024 (config.c:176-179).

Word order: arrays here are native u128 -- numpy uint64 (n, 2) with column 0
the low word, or torch int64 (n, 2) on the GPU.  The reference's AVX code and
files use hi/lo order; the converters are the library's
(bntt_hilo_to_u128[_dev], bntt_u128_to_hilo[_dev]).

Generated inputs (library kernels, include/bntt.h): root-of-unity twiddle
tables (bntt_twiddles[_dev]) and seeded uniform residues
(bntt_uniform[_dev]).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (BNTT_TW_CYCLIC, BNTT_TW_CYCLIC_INV, BNTT_TW_NEGACYCLIC,  # noqa: F401
                   BNTT_TW_NEGACYCLIC_INV, U128)
from .ops import _dev_ptr, _host_ptr, _is_torch, _nelem, _stream

KINDS = {"cyclic": BNTT_TW_CYCLIC, "negacyclic": BNTT_TW_NEGACYCLIC, "cyclic_inv": BNTT_TW_CYCLIC_INV,
         "negacyclic_inv": BNTT_TW_NEGACYCLIC_INV}


# ------------------------------------------------------------------ text files
def read_text(path: str, n: int | None = None) -> np.ndarray:
    """Residues of a reference-format text file, native layout."""
    with open(path) as f:
        words = np.array([int(w) for w in f.read().split()], dtype=np.uint64)
    if words.size % 2:
        raise ValueError(f"{path}: odd number of words")
    arr = words.reshape(-1, 2)
    if n is not None:
        if arr.shape[0] < n:
            raise ValueError(f"{path}: {arr.shape[0]} residues, {n} needed")
        arr = arr[:n]
    return hilo_to_u128(arr)


def write_text(path: str, a: np.ndarray, final_newline: bool = False) -> None:
    """Write residues (native layout) in the reference's text format."""
    hl = u128_to_hilo(np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2))
    with open(path, "w") as f:
        f.write("\n".join(str(int(w)) for w in hl.reshape(-1)) + ("\n" if final_newline else ""))


def tile(a1024: np.ndarray, n: int) -> np.ndarray:
    """The reference's input for N > 1024: the 1024-point array repeated."""
    a1024 = np.asarray(a1024).reshape(-1, 2)
    if n <= a1024.shape[0]:
        return np.ascontiguousarray(a1024[:n])
    if n % a1024.shape[0]:
        raise ValueError("N must be a multiple of the stored length")
    return np.ascontiguousarray(np.tile(a1024, (n // a1024.shape[0], 1)))


def load_ntt_set(data_dir: str, n: int) -> dict:
    """x, twiddles and (N <= 1024) the expected output, as the reference's
    NTT drivers load them."""
    src = n if n <= 1024 else 1024
    out = {"x": tile(read_text(os.path.join(data_dir, f"x_{src}.txt"), src), n),
           "tw": tile(read_text(os.path.join(data_dir, f"twd_{src}.txt"), src), n),
           "ver": None}
    if n <= 1024:
        out["ver"] = read_text(os.path.join(data_dir, f"ver_{n}.txt"), n)
    return out


def load_blas_set(data_dir: str) -> dict:
    """Inputs and expected outputs of the reference's BLAS driver: a = twd_1024,
    b = x_1024 (blas_scalar.c:618-749)."""
    out = {"a": read_text(os.path.join(data_dir, "twd_1024.txt"), 1024),
           "b": read_text(os.path.join(data_dir, "x_1024.txt"), 1024)}
    for op in ("vvadd", "vvsub", "vvmul", "axpy"):
        out[op] = read_text(os.path.join(data_dir, f"{op}_1024.txt"), 1024)
    return out


# ------------------------------------------------------------ word order
def hilo_to_u128(hl, out=None, stream=None):
    """hi/lo word order -> native u128 (numpy on the host, torch on the GPU)."""
    return _convert("bntt_hilo_to_u128", hl, out, stream)


def u128_to_hilo(a, out=None, stream=None):
    """native u128 -> hi/lo word order."""
    return _convert("bntt_u128_to_hilo", a, out, stream)


def _convert(name, src, out, stream):
    n = _nelem(src)
    if out is not None and _nelem(out) != n:
        raise ValueError("out must hold as many residues as the input")
    if _is_torch(src):
        if out is None:
            import torch

            out = torch.empty_like(src)
        _lib.call(name + "_dev", _dev_ptr(out), _dev_ptr(src), n, _stream(stream))
        return out
    src = np.ascontiguousarray(src, dtype=np.uint64).reshape(-1, 2)
    if out is None:
        out = np.empty_like(src)
    _lib.call(name, _host_ptr(out, out=True), _host_ptr(src), n)
    return out


# ------------------------------------------------------------ generated inputs
def twiddles(log2n: int, q: int, kind: str = "negacyclic", device: bool = True, stream=None):
    """Root-of-unity table of 2^log2n entries (kinds: cyclic, negacyclic,
    cyclic_inv, negacyclic_inv), for a prime q < 2^122 with enough factors of
    two in q - 1."""
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {sorted(KINDS)}")
    k = KINDS[kind]
    qq = U128.of(q)
    n = 1 << log2n
    if device:
        import torch

        t = torch.empty((n, 2), dtype=torch.int64, device="cuda")
        _lib.call("bntt_twiddles_dev", _dev_ptr(t), log2n, ctypes.byref(qq), k, _stream(stream))
        return t
    a = np.empty((n, 2), dtype=np.uint64)
    _lib.call("bntt_twiddles", _host_ptr(a), log2n, ctypes.byref(qq), k)
    return a


def uniform(n: int, q: int, seed: int, device: bool = True, stream=None):
    """n residues uniform in [0, q) from the library's counter-based splitmix64
    generator (the same values on the host and the GPU)."""
    qq = U128.of(q)
    if device:
        import torch

        t = torch.empty((n, 2), dtype=torch.int64, device="cuda")
        _lib.call("bntt_uniform_dev", _dev_ptr(t), n, ctypes.byref(qq), seed, _stream(stream))
        return t
    a = np.empty((n, 2), dtype=np.uint64)
    _lib.call("bntt_uniform", _host_ptr(a), n, ctypes.byref(qq), seed)
    return a


# ------------------------------------------------------------ scalar primitives
def addmod128(a: int, b: int, m: int) -> int:
    """ntt_scalar.c:49-79, bit-identical (host)."""
    return _lib.load().bntt_addmod128(U128.of(a), U128.of(b), U128.of(m)).value


def submod128(a: int, b: int, m: int) -> int:
    """ntt_scalar.c:81-104, bit-identical (host)."""
    return _lib.load().bntt_submod128(U128.of(a), U128.of(b), U128.of(m)).value


def mulmod128(a: int, b: int, m: int, mu: int) -> int:
    """ntt_scalar.c:173-199 (Barrett, MBITS = 124), bit-identical (host)."""
    return _lib.load().bntt_mulmod128(U128.of(a), U128.of(b), U128.of(m), U128.of(mu)).value
