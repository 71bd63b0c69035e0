"""Reference-shaped host interface over the C ABI.

The functions mirror the reference's kernels (files of the reference's src/):
    ntt(Y, X, modulus, mu, twiddles, as0, as1)   ntt_scalar.c:582-627
    vector_add(c, a, b, m)                       blas_scalar.c:578-584
    vector_sub(c, a, b, m)                       blas_scalar.c:586-592
    vector_mul(c, a, b, m, mu)                   blas_scalar.c:594-600
    axpy(c, a, b, x, m, mu)                      blas_scalar.c:602-608
with the same argument order and meaning: outputs first, written in place,
modulus/mu/scalars as Python ints.  Sizes come from the arrays instead of
compile-time macros.  Arrays are either torch tensors on the GPU (the device
entry points; work is enqueued on torch's current stream) or numpy arrays in
host memory (the host entry points; synchronous).  Either way each element
is 16 bytes in unsigned __int128 layout: numpy uint64 (n, 2) or torch int64
(n, 2).  Errors raise BnttError; nothing runs on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import U128


def _is_torch(a) -> bool:
    return type(a) is not np.ndarray and type(a).__module__.startswith("torch")


def _nelem(a) -> int:
    nbytes = a.nbytes  # numpy arrays and torch tensors alike
    if nbytes % 16:
        raise ValueError("residue arrays hold 16 bytes per element")
    return nbytes // 16


def _dev_ptr(a) -> int:
    if not a.is_cuda:
        raise ValueError("torch tensors must be on the GPU")
    if not a.is_contiguous():
        raise ValueError("tensors must be contiguous")
    return a.data_ptr()


def _host_ptr(a: np.ndarray, out: bool = False) -> int:
    """Address of a host array's data; out: the library writes it."""
    if not isinstance(a, np.ndarray):
        raise ValueError("host arrays must be C-contiguous numpy arrays")
    fl = a.flags
    if not fl.c_contiguous:
        raise ValueError("host arrays must be C-contiguous numpy arrays")
    if out and not fl.writeable:
        raise ValueError("output arrays must be writeable")
    if fl.writeable and a.nbytes:
        # the buffer protocol gives the address faster than a.ctypes
        return ctypes.addressof(_c_char.from_buffer(a))
    return a.ctypes.data


_c_char = ctypes.c_char


_U128_CACHE: dict = {}


def _u128(v: int) -> U128:
    """The (read-only) U128 of a modulus or scalar, reused across calls."""
    c = _U128_CACHE.get(v)
    if c is None:
        if len(_U128_CACHE) > 64:
            _U128_CACHE.clear()
        c = _U128_CACHE[v] = U128.of(v)
    return c


_raw_stream = None


def _current_raw_stream(device_index: int) -> int:
    """torch's current stream on the device, as a hipStream_t handle (torch's
    own raw accessor where it has one: building a Stream object per call costs
    several microseconds)."""
    global _raw_stream
    if _raw_stream is None:
        import torch

        f = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        _raw_stream = f if f is not None else (lambda d: torch.cuda.current_stream(d).cuda_stream)
    return _raw_stream(device_index)


def _stream(stream=None, like=None):
    """The stream argument of a device entry point: `stream` (a torch stream
    object or a raw hipStream_t handle), else torch's current stream on the
    device of tensor `like` (or the current device)."""
    if stream is not None:
        return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))
    if like is not None:
        return ctypes.c_void_p(_current_raw_stream(like.get_device()))
    import torch

    return ctypes.c_void_p(_current_raw_stream(torch.cuda.current_device()))


def _mode(*arrays) -> str:
    kinds = {_is_torch(a) for a in arrays}
    if len(kinds) != 1:
        raise ValueError("mixing torch tensors and numpy arrays")
    return "dev" if kinds.pop() else "host"


def _same_len(*arrays) -> int:
    ns = {_nelem(a) for a in arrays}
    if len(ns) != 1:
        raise ValueError("arrays differ in length")
    return ns.pop()


def _blas(name, c, a, b, extra_before_n, m, mu, stream):
    n = _same_len(c, a, b)
    mm = _u128(m)
    args_mu = [ctypes.byref(_u128(mu))] if mu is not None else []
    if _mode(c, a, b) == "dev":
        ptrs = [_dev_ptr(c), _dev_ptr(a), _dev_ptr(b)]
        _lib.call(f"bntt_{name}_dev", *ptrs, *extra_before_n, n, ctypes.byref(mm), *args_mu, _stream(stream, c))
    else:
        ptrs = [_host_ptr(c, out=True), _host_ptr(a), _host_ptr(b)]
        _lib.call(f"bntt_{name}", *ptrs, *extra_before_n, n, ctypes.byref(mm), *args_mu)


def vector_add(c, a, b, m: int, stream=None) -> None:
    """c = a + b mod m (blas_scalar.c:578-584)."""
    _blas("vvadd", c, a, b, [], m, None, stream)


def vector_sub(c, a, b, m: int, stream=None) -> None:
    """c = a - b mod m (blas_scalar.c:586-592)."""
    _blas("vvsub", c, a, b, [], m, None, stream)


def vector_mul(c, a, b, m: int, mu: int, stream=None) -> None:
    """c = a * b mod m (blas_scalar.c:594-600)."""
    _blas("vvmul", c, a, b, [], m, mu, stream)


def axpy(c, a, b, x: int, m: int, mu: int, stream=None) -> None:
    """c = a + x * b mod m (blas_scalar.c:602-608)."""
    xs = U128.of(x)
    _blas("axpy", c, a, b, [ctypes.byref(xs)], m, mu, stream)


_TMP_CACHE: dict = {}


def tmp_elements(log2n: int, count: int = 1) -> int:
    """Residues of scratch the device NTT uses: 0 for log2n <= 12, except a
    lone transform of 2^11 .. 2^13 points (2^log2n + 16 residues: with it the
    transform splits over several workgroups; optional up to 2^12), and
    count * 2^log2n above 2^12 points otherwise."""
    key = (log2n, count)
    v = _TMP_CACHE.get(key)
    if v is None:
        if len(_TMP_CACHE) > 256:
            _TMP_CACHE.clear()
        v = _TMP_CACHE[key] = _lib.load().bntt_ntt_tmp_bytes(log2n, count) // 16
    return v


_fwd_host = None


def _ntt_host(Y, X, tw, modulus, mu) -> None:
    """ntt() on numpy arrays: one synchronous host-array call per transform
    (the same checks as the general path, in fewer Python steps)."""
    global _fwd_host
    nb = tw.nbytes
    total = X.nbytes
    if nb % 16 or total % 16:
        raise ValueError("residue arrays hold 16 bytes per element")
    if Y.nbytes != total:
        raise ValueError("arrays differ in length")
    n = nb >> 4
    log2n = n.bit_length() - 1
    if n != 1 << log2n:
        raise ValueError("N must be a power of two")
    if (total >> 4) % n:
        raise ValueError("X must hold a whole number of transforms")
    count = (total >> 4) // n
    py, px, pt = _host_ptr(Y, out=True), _host_ptr(X), _host_ptr(tw)
    if _fwd_host is None:
        _fwd_host = _lib.load().bntt_ntt_fwd
    mm, mmu = ctypes.byref(_u128(modulus)), ctypes.byref(_u128(mu))
    step = n << 4
    for i in range(count):
        rc = _fwd_host(py, px, pt, log2n, mm, mmu)
        if rc:
            _lib.check(rc)
        py += step
        px += step


def ntt(Y, X, modulus: int, mu: int, twiddles, as0=None, as1=None, stream=None) -> None:
    """Forward NTT (ntt_scalar.c:582-627): Y = NTT(X) with the twiddle table
    `twiddles` (N entries).  N = len(twiddles); X and Y hold a whole number of
    transforms of N points (one, as in the reference, or a uniform batch).
    as0/as1 are the reference's scratch arguments: on the device path as0
    serves as the scratch when the launch uses one (one transform of 2^11 ..
    2^13 points, which then splits over several workgroups: len >= N + 16,
    else a scratch of that size is allocated; several transforms of more than
    4096 points: len >= len(X)); as1 is accepted and unused."""
    del as1
    if type(Y) is np.ndarray and type(X) is np.ndarray and type(twiddles) is np.ndarray:
        _ntt_host(Y, X, twiddles, modulus, mu)
        return
    n = _nelem(twiddles)
    log2n = n.bit_length() - 1
    if n != 1 << log2n:
        raise ValueError("N must be a power of two")
    total = _same_len(Y, X)
    if total % n:
        raise ValueError("X must hold a whole number of transforms")
    count = total // n
    mm, mmu = _u128(modulus), _u128(mu)
    if _mode(Y, X, twiddles) == "dev":
        need = tmp_elements(log2n, count)
        tmp = 0
        own = None
        if need:
            if as0 is not None and _nelem(as0) < need:
                if count != 1 or log2n > 13:
                    raise ValueError(f"as0 must hold {need} residues for this N")
                # too small for the split of one transform (N + 16, where the
                # reference's as0 holds N): use our own
                as0 = None
            if as0 is None:
                import torch

                as0 = own = torch.empty((need, 2), dtype=torch.int64, device=X.device)
            tmp = _dev_ptr(as0)
        py, px, pt = _dev_ptr(Y), _dev_ptr(X), _dev_ptr(twiddles)
        s = _stream(stream, X)
        _lib.call("bntt_ntt_fwd_dev", py, px, pt, log2n, count, ctypes.byref(mm), ctypes.byref(mmu), tmp, s)
        if own is not None and stream is not None:
            # the scratch was allocated on torch's current stream but is
            # written on `stream`: keep the allocator from reusing it before
            # the launch's kernels have run
            import torch

            own.record_stream(torch.cuda.ExternalStream(s.value or 0, device=X.device))
    else:
        py, px, pt = _host_ptr(Y, out=True), _host_ptr(X), _host_ptr(twiddles)
        for i in range(count):
            off = i * n * 16
            _lib.call("bntt_ntt_fwd", py + off, px + off, pt, log2n, ctypes.byref(mm), ctypes.byref(mmu))
