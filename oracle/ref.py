"""Driver for the reference's own kernels compiled by build_ref.sh into _ref/.

TEST INFRASTRUCTURE ONLY (second checker and CPU baseline).  Each object in
_ref/ holds one reference source file (ntt_{scalar,avx2,avx512}.c at a fixed
`stages`, or blas_{scalar,avx2,avx512}.c) plus the pointer-only wrappers of
ref_glue.c.  Objects are loaded RTLD_LOCAL, so the identically named reference
symbols of different files never meet.

Layouts: scalar kernels take native u128 arrays ((n, 2) uint64, low word
first); the AVX kernels take the reference's hi/lo-u64 layout (high word
first), converted here outside any timed region.
"""
from __future__ import annotations

import ctypes
import os
import threading
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_P = ctypes.c_void_p
_LOADED: dict[str, ctypes.CDLL] = {}
_LOCK = threading.Lock()

BLAS_OPS = {"vvadd": 0, "vvsub": 1, "vvmul": 2, "axpy": 3}


def cpu_flags() -> set[str]:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def best_isa() -> str:
    fl = cpu_flags()
    if {"avx512f", "avx512dq", "avx512bw", "avx512vl"} <= fl:
        return "avx512"
    if "avx2" in fl and "bmi2" in fl:
        return "avx2"
    return "scalar"


# The ISA extensions -march=znver4 may use (the reference's AMD build,
# Makefile:4-5): the "amd" objects run only where all are present.
ZNVER4_FLAGS = {"avx2", "bmi2", "adx", "avx512f", "avx512dq", "avx512cd", "avx512bw", "avx512vl", "avx512ifma",
                "avx512vbmi", "avx512_vbmi2", "avx512_vnni", "avx512_bitalg", "avx512_vpopcntdq", "avx512_bf16",
                "gfni", "vaes", "vpclmulqdq"}

# compiler and flags of each flavour (oracle/build_ref.sh)
FLAVOUR_FLAGS = {
    "amd": "clang (ROCm 7.2, LLVM 22) -march=znver4 -O3 -w: the reference's AMD configuration, Makefile:4-5",
    "portable": "gcc 11 -O3 -w -march=x86-64-v3 (scalar, AVX2) / x86-64-v4 (AVX-512)",
}


def flavour() -> str:
    """'amd' where the host has znver4's ISA and those objects were built,
    else 'portable'."""
    if ZNVER4_FLAGS <= cpu_flags() and os.path.exists(os.path.join(REF_DIR, "ref_blas_scalar_amd.so")):
        return "amd"
    return "portable"


def _sfx(fl: str | None) -> str:
    return "_amd" if (fl or flavour()) == "amd" else ""


def cpu_threads() -> int:
    """The cores this process may use: its affinity set, capped by a cgroup
    CPU quota when one is set (cpu.max)."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def host_info() -> dict:
    """What the CPU baseline ran on: core counts, model and the flags of the
    reference objects used."""
    model = cpu_model()
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    fl = cpu_flags()
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "threads_used": cpu_threads(),
            "model": model, "isa": sorted(f for f in fl if f.startswith(("avx", "amx")) or f in ZNVER4_FLAGS),
            "flavour": flavour(), "compiler_flags": FLAVOUR_FLAGS[flavour()]}
    # a virtualised host may report only the brand as its model name: the
    # family / model / stepping numbers and the clock identify the generation
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                names = {"cpu family": "cpu_family", "model": "cpu_model_number", "stepping": "cpu_stepping",
                         "cpu MHz": "cpu_mhz", "cache size": "cpu_cache"}
                if key in names and names[key] not in info:
                    info[names[key]] = val
                if not line.strip():
                    break
    except OSError:
        pass
    return info


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _lib(name: str) -> ctypes.CDLL:
    with _LOCK:
        if name not in _LOADED:
            path = os.path.join(REF_DIR, name)
            if not os.path.exists(path):
                raise FileNotFoundError(path)
            lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
            lib.bntt_ref_init.restype = ctypes.c_int
            lib.bntt_ref_init()
            _LOADED[name] = lib
        return _LOADED[name]


def available(name: str = "ref_blas_scalar.so") -> bool:
    return os.path.exists(os.path.join(REF_DIR, name))


def _u128(v: int) -> np.ndarray:
    return np.array([v & (2**64 - 1), v >> 64], dtype=np.uint64)


def native_to_hilo(a: np.ndarray) -> np.ndarray:
    out = np.empty(a.size, dtype=np.uint64)
    out[0::2] = a.reshape(-1, 2)[:, 1]
    out[1::2] = a.reshape(-1, 2)[:, 0]
    return out


def hilo_to_native(w: np.ndarray) -> np.ndarray:
    w = w.reshape(-1, 2)
    return np.ascontiguousarray(w[:, ::-1])


def _aligned(nwords: int) -> np.ndarray:
    buf = np.empty(nwords + 8, dtype=np.uint64)
    off = (-buf.ctypes.data // 8) % 8
    return buf[off:off + nwords]


def ntt(x: np.ndarray, tw: np.ndarray, m: int, mu: int, isa: str = "scalar", iters: int = 0,
        fl: str | None = "portable"):
    """Run the reference ntt at N = len(tw).  Returns (y, ns_per_call): the
    mean of the last half of `iters` timed calls after one warm call (the
    reference's protocol, ntt_scalar.c:670-683, with iters = 100).  fl: the
    build flavour ('portable', 'amd', or None for the host's best)."""
    n = tw.reshape(-1, 2).shape[0]
    log2n = n.bit_length() - 1
    lib = _lib(f"ref_ntt_{isa}{_sfx(fl)}_L{log2n}.so")
    lib.bntt_ref_ntt.restype = ctypes.c_double
    lib.bntt_ref_ntt.argtypes = [_P, _P, _P, _P, _P, ctypes.c_int]
    mm, mmu = _u128(m), _u128(mu)
    if isa == "scalar":
        xx = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2).copy()
        tt = np.ascontiguousarray(tw, dtype=np.uint64).reshape(-1, 2).copy()
        y = np.empty_like(xx)
        dt = lib.bntt_ref_ntt(y.ctypes.data, xx.ctypes.data, mm.ctypes.data, mmu.ctypes.data,
                              tt.ctypes.data, iters)
        return y, dt
    xx = _aligned(2 * n)
    xx[:] = native_to_hilo(np.asarray(x, dtype=np.uint64))
    tt = _aligned(2 * n)
    tt[:] = native_to_hilo(np.asarray(tw, dtype=np.uint64))
    yy = _aligned(2 * n)
    dt = lib.bntt_ref_ntt(yy.ctypes.data, xx.ctypes.data, mm.ctypes.data, mmu.ctypes.data,
                          tt.ctypes.data, iters)
    return hilo_to_native(yy), dt


class BlasOperands:
    """Operands laid out once for one reference BLAS variant (untimed)."""

    def __init__(self, a: np.ndarray, b: np.ndarray, isa: str, fl: str | None = "portable"):
        self.lib_name = f"ref_blas_{isa}{_sfx(fl)}.so"
        a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
        b = np.asarray(b, dtype=np.uint64).reshape(-1, 2)
        n = a.shape[0]
        assert n % 1024 == 0, "the reference BLAS kernels work on 1024-element blocks"
        self.n, self.isa = n, isa
        if isa == "scalar":
            self.a = np.ascontiguousarray(a).copy()
            self.b = np.ascontiguousarray(b).copy()
            self.c = np.empty_like(self.a)
        else:
            self.a = _aligned(2 * n)
            self.a[:] = native_to_hilo(a)
            self.b = _aligned(2 * n)
            self.b[:] = native_to_hilo(b)
            self.c = _aligned(2 * n)

    def result(self) -> np.ndarray:
        return self.c.copy() if self.isa == "scalar" else hilo_to_native(self.c)


def blas(op: str, ops: BlasOperands, m: int, mu: int, x: int = 0, threads: int = 1) -> float:
    """Run the reference kernel `op` over all 1024-blocks of `ops`, split into
    contiguous block ranges over `threads` threads (ctypes releases the GIL).
    Returns wall-clock seconds."""
    lib = _lib(ops.lib_name)
    lib.bntt_ref_blas.restype = ctypes.c_double
    lib.bntt_ref_blas.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_long, _P, _P, _P]
    mm, mmu, xx = _u128(m), _u128(mu), _u128(x)
    nchunks = ops.n // 1024
    words = 2  # u64 words per element in either layout
    threads = max(1, min(threads, nchunks))
    bounds = [nchunks * t // threads for t in range(threads + 1)]

    def run(lo: int, hi: int):
        off = lo * 1024 * words * 8
        lib.bntt_ref_blas(BLAS_OPS[op], ops.c.ctypes.data + off, ops.a.ctypes.data + off,
                          ops.b.ctypes.data + off, hi - lo, mm.ctypes.data, mmu.ctypes.data,
                          xx.ctypes.data)

    t0 = time.perf_counter()
    if threads == 1:
        run(0, nchunks)
    else:
        ths = [threading.Thread(target=run, args=(bounds[t], bounds[t + 1])) for t in range(threads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    return time.perf_counter() - t0


def blas_ns_per_element(op: str, ops: BlasOperands, m: int, mu: int, x: int = 0, iters: int = 100) -> float:
    """The reference's BLAS protocol on the first 1024 elements of `ops`
    (blas_scalar.c:660-721): `iters` calls, ns per element of the mean of the
    last half, on the calling thread."""
    lib = _lib(ops.lib_name)
    lib.bntt_ref_blas_repeat.restype = ctypes.c_double
    lib.bntt_ref_blas_repeat.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_int, _P, _P, _P]
    mm, mmu, xx = _u128(m), _u128(mu), _u128(x)
    ns = lib.bntt_ref_blas_repeat(BLAS_OPS[op], ops.c.ctypes.data, ops.a.ctypes.data, ops.b.ctypes.data, iters,
                                  mm.ctypes.data, mmu.ctypes.data, xx.ctypes.data)
    return ns / 1024


def scalar_primitive(name: str, a: int, b: int, m: int, mu: int = 0) -> int:
    """addmod128 / submod128 / mulmod128 of blas_scalar.c (copies of
    ntt_scalar.c:49-199)."""
    lib = _lib("ref_blas_scalar.so")
    fn = getattr(lib, f"bntt_ref_{name}")
    r = np.zeros(2, np.uint64)
    aa, bb, mm, mmu = _u128(a), _u128(b), _u128(m), _u128(mu)
    if name == "mulmod128":
        fn.argtypes = [_P] * 5
        fn(r.ctypes.data, aa.ctypes.data, bb.ctypes.data, mm.ctypes.data, mmu.ctypes.data)
    else:
        fn.argtypes = [_P] * 4
        fn(r.ctypes.data, aa.ctypes.data, bb.ctypes.data, mm.ctypes.data)
    return int(r[0]) | (int(r[1]) << 64)
