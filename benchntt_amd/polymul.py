"""Negacyclic polynomial multiplication with continuous batching (C ABI:
bntt_pm_* in include/bntt.h).

    srv = PolymulServer(q=config.Q_116, hbm_budget=1 << 30)
    srv.multiply([(a, b, c), ...])              # one batch, returns when done
    done_ns, stats = srv.run_trace(reqs, arrival_ns)  # DROPPED: not admitted in time
    srv.forward(y, x, n); srv.inverse(x, y, n)  # the pipeline's two transforms

a, b, c are device tensors of N canonical residues (torch int64 (N, 2)); the
product c = a * b mod (x^N + 1, q).  forward / inverse take count transforms of
n points back to back: the negacyclic NTT with outputs in bit-reversed order,
and its inverse.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BNTT_PM_DROPPED as DROPPED  # noqa: F401  (done_ns of a dropped request)
from ._lib import PmConfig, PmRequest, PmStats, U128
from .ops import _dev_ptr, _nelem


def _ready():
    # the server runs on streams of its own that do not wait for other streams
    import torch

    torch.cuda.current_stream().synchronize()


def _requests(reqs):
    arr = (PmRequest * max(1, len(reqs)))()
    for i, (a, b, c) in enumerate(reqs):
        n = _nelem(c)
        l = n.bit_length() - 1
        if n != 1 << l or _nelem(a) != n or _nelem(b) != n:
            raise ValueError("a, b and c must hold the same power-of-two number of residues")
        arr[i] = PmRequest(_dev_ptr(a), _dev_ptr(b), _dev_ptr(c), l, 0)
    return arr


class PolymulServer:
    def __init__(self, q: int, hbm_budget: int = 1 << 30, max_inflight: int = 4, max_batch: int = 64,
                 min_log2n: int = 1, max_log2n: int = 16, admit_timeout_ns: int = 0,
                 max_batch_butterflies: int = 0):
        """admit_timeout_ns: run_trace drops a request not admitted within
        this time of its arrival (0: never).  max_batch_butterflies: run_trace
        closes a batch before a request that would take its work (3 * N/2 *
        log2 N butterflies per request) above this (0: no limit)."""
        cfg = PmConfig(hbm_budget, max_inflight, max_batch, min_log2n, max_log2n, admit_timeout_ns,
                       max_batch_butterflies)
        h = ctypes.c_void_p()
        qq = U128.of(q)
        _lib.call("bntt_pm_server_create", ctypes.byref(h), ctypes.byref(qq), ctypes.byref(cfg))
        self._h = h
        self.max_batch = max_batch

    def multiply(self, reqs) -> None:
        reqs = list(reqs)
        _ready()
        for i in range(0, len(reqs), self.max_batch):
            chunk = reqs[i:i + self.max_batch]
            _lib.call("bntt_pm_multiply", self._h, _requests(chunk), len(chunk))

    def _transform(self, direction, y, x, n):
        total = _nelem(x)
        l = n.bit_length() - 1
        if n != 1 << l or total % n or _nelem(y) != total:
            raise ValueError("x and y must hold the same whole number of n-point transforms, n a power of two")
        _ready()
        _lib.call("bntt_pm_transform", self._h, direction, _dev_ptr(y), _dev_ptr(x), l, total // n)

    def forward(self, y, x, n: int) -> None:
        """y = the negacyclic NTT of each n-point block of x, outputs of a
        block in bit-reversed order (y may be x)."""
        self._transform(_lib.BNTT_PM_FORWARD, y, x, n)

    def inverse(self, y, x, n: int) -> None:
        """The inverse of forward: x in forward's output order, y the
        coefficients (y may be x)."""
        self._transform(_lib.BNTT_PM_INVERSE, y, x, n)

    def run_trace(self, reqs, arrival_ns):
        reqs = list(reqs)
        arrival = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        if arrival.shape[0] != len(reqs):
            raise ValueError("one arrival time per request")
        _ready()
        done = np.zeros(len(reqs), dtype=np.uint64)
        st = PmStats()
        _lib.call("bntt_pm_run_trace", self._h, _requests(reqs),
                  arrival.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(reqs),
                  done.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(st))
        return done, {"requests": st.requests, "batches": st.batches,
                      "max_inflight_requests": st.max_inflight_requests, "admission_waits": st.admission_waits,
                      "dropped": st.dropped, "max_batch_ns": st.max_batch_ns,
                      "max_loop_gap_ns": st.max_loop_gap_ns}

    def close(self) -> None:
        if self._h:
            _lib.call("bntt_pm_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
