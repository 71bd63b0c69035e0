"""ctypes binding of libbntt.so (the C ABI of include/bntt.h).

The library is built in-tree by ``python benchntt_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for anything here.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbntt.so")

BNTT_OK = 0
BNTT_EINVAL = 1  # bad size, pointer or alignment
BNTT_EHIP = 2  # HIP runtime error
BNTT_ENODEV = 3  # no gfx950 device
BNTT_ENOMEM = 4  # device allocation failed


class BnttError(RuntimeError):
    """A bntt_* entry point returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"bntt error {code}: {msg}")
        self.code = code


class U128(ctypes.Structure):
    """bntt_u128: a 128-bit value in native unsigned __int128 layout."""

    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]

    @classmethod
    def of(cls, v: int) -> "U128":
        if v < 0 or v >= 1 << 128:
            raise ValueError("value does not fit in 128 bits")
        return cls(v & (2**64 - 1), v >> 64)

    @property
    def value(self) -> int:
        return self.lo | (self.hi << 64)


class NttTask(ctypes.Structure):
    """bntt_ntt_task."""

    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("tw", ctypes.c_void_p),
                ("log2n", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MulTask(ctypes.Structure):
    """bntt_mul_task."""

    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p), ("n", ctypes.c_uint64)]


class TuneEntry(ctypes.Structure):
    """bntt_tune_entry."""

    _fields_ = [("log2n", ctypes.c_int32), ("elog", ctypes.c_int32), ("passes", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class BatchInfo(ctypes.Structure):
    """bntt_batch_info."""

    _fields_ = [("levels", ctypes.c_int32), ("reserved", ctypes.c_int32), ("items", ctypes.c_uint64),
                ("tmp_bytes", ctypes.c_uint64), ("level_items", ctypes.c_uint64 * 8),
                ("level_grid", ctypes.c_int32 * 8), ("level_lds_bytes", ctypes.c_uint32 * 8)]


class PmRequest(ctypes.Structure):
    """bntt_pm_request."""

    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("log2n", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PmConfig(ctypes.Structure):
    """bntt_pm_config."""

    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("min_log2n", ctypes.c_int32), ("max_log2n", ctypes.c_int32),
                ("admit_timeout_ns", ctypes.c_uint64), ("max_batch_butterflies", ctypes.c_uint64)]


class PmStats(ctypes.Structure):
    """bntt_pm_stats."""

    _fields_ = [("requests", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("max_inflight_requests", ctypes.c_uint64), ("admission_waits", ctypes.c_uint64),
                ("dropped", ctypes.c_uint64), ("max_batch_ns", ctypes.c_uint64),
                ("max_loop_gap_ns", ctypes.c_uint64)]


# done_ns of a request that bntt_pm_run_trace dropped
BNTT_PM_DROPPED = 2**64 - 1


BNTT_PLAN_GRAPH = 1
BNTT_PLAN_TW_FIXED = 8

# residues of the fixed scratch of bntt_ntt_fwd_dev_noalloc (BNTT_NOALLOC_MAX_RESIDUES)
BNTT_NOALLOC_MAX_RESIDUES = 1 << 20

_P = ctypes.c_void_p
_PU = ctypes.POINTER(U128)
_SZ = ctypes.c_size_t

# name -> argtypes (all return int status, except where noted)
_SIGS = {
    "bntt_version": [],
    "bntt_init": [],
    "bntt_vvadd_dev": [_P, _P, _P, _SZ, _PU, _P],
    "bntt_vvsub_dev": [_P, _P, _P, _SZ, _PU, _P],
    "bntt_vvmul_dev": [_P, _P, _P, _SZ, _PU, _PU, _P],
    "bntt_axpy_dev": [_P, _P, _P, _PU, _SZ, _PU, _PU, _P],
    "bntt_ntt_fwd_dev": [_P, _P, _P, ctypes.c_int, _SZ, _PU, _PU, _P, _P],
    "bntt_ntt_fwd_dev_noalloc": [_P, _P, _P, ctypes.c_int, _SZ, _PU, _PU, _P],
    "bntt_vvadd": [_P, _P, _P, _SZ, _PU],
    "bntt_vvsub": [_P, _P, _P, _SZ, _PU],
    "bntt_vvmul": [_P, _P, _P, _SZ, _PU, _PU],
    "bntt_axpy": [_P, _P, _P, _PU, _SZ, _PU, _PU],
    "bntt_ntt_fwd": [_P, _P, _P, ctypes.c_int, _PU, _PU],
    "bntt_batch_plan_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(NttTask), _SZ,
                               ctypes.POINTER(MulTask), _SZ, _PU, _PU, ctypes.POINTER(TuneEntry), _SZ,
                               ctypes.c_int],
    "bntt_batch_plan_execute": [_P, _P],
    "bntt_batch_plan_info": [_P, ctypes.POINTER(BatchInfo)],
    "bntt_batch_plan_destroy": [_P],
    "bntt_ntt_launch_each_dev": [ctypes.POINTER(NttTask), _SZ, _PU, _PU, _P, _P],
    "bntt_pm_server_create": [ctypes.POINTER(ctypes.c_void_p), _PU, ctypes.POINTER(PmConfig)],
    "bntt_pm_server_destroy": [_P],
    "bntt_pm_multiply": [_P, ctypes.POINTER(PmRequest), _SZ],
    "bntt_pm_transform": [_P, ctypes.c_int, _P, _P, ctypes.c_int, _SZ],
    "bntt_pm_run_trace": [_P, ctypes.POINTER(PmRequest), ctypes.POINTER(ctypes.c_uint64), _SZ,
                          ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(PmStats)],
    "bntt_hilo_to_u128": [_P, _P, _SZ],
    "bntt_u128_to_hilo": [_P, _P, _SZ],
    "bntt_hilo_to_u128_dev": [_P, _P, _SZ, _P],
    "bntt_u128_to_hilo_dev": [_P, _P, _SZ, _P],
    "bntt_twiddles_dev": [_P, ctypes.c_int, _PU, ctypes.c_int, _P],
    "bntt_twiddles": [_P, ctypes.c_int, _PU, ctypes.c_int],
    "bntt_uniform_dev": [_P, _SZ, _PU, ctypes.c_uint64, _P],
    "bntt_uniform": [_P, _SZ, _PU, ctypes.c_uint64],
    "bntt_probe_valu_rate": [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
}
# entry points that do not return a status
_OTHER = {
    "bntt_last_error": ([], ctypes.c_char_p),
    "bntt_ntt_tmp_bytes": ([ctypes.c_int, _SZ], _SZ),
    "bntt_addmod128": ([U128, U128, U128], U128),
    "bntt_submod128": ([U128, U128, U128], U128),
    "bntt_mulmod128": ([U128, U128, U128, U128], U128),
}

BNTT_TW_CYCLIC, BNTT_TW_NEGACYCLIC, BNTT_TW_CYCLIC_INV, BNTT_TW_NEGACYCLIC_INV = 0, 1, 2, 3
BNTT_PM_FORWARD, BNTT_PM_INVERSE = 0, 1

EXPORTS = tuple(_SIGS) + tuple(_OTHER)

_lib: ctypes.CDLL | None = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libbntt.so once and attach prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: build the HIP library first (python benchntt_amd/build.py)"
        )
    # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.  Importing
    # it first makes libbntt.so bind to those same copies (same SONAMEs), so the
    # process has one HIP runtime and torch's device pointers and streams are
    # valid in libbntt.  Loaded the other way round, the process would hold two
    # HSA runtimes and the second one sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    missing = [n for n in EXPORTS if not hasattr(lib, n)]
    if missing:
        raise ImportError(f"{p} lacks {', '.join(missing)}: rebuild it (python benchntt_amd/build.py)")
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    for name, (args, res) in _OTHER.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != BNTT_OK:
        msg = load().bntt_last_error()
        raise BnttError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
