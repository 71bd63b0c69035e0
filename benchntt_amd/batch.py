"""Uneven batches: many forward NTTs of different sizes, and pointwise
products of their outputs, in one persistent work-queue kernel (C ABI:
bntt_batch_plan_* in include/bntt.h).

    plan = BatchPlan(tasks=[(x, y, tw), ...], muls=[(a, b, c), ...],
                     modulus=q, mu=mu, tune=load_tune_table())
    plan.execute()          # enqueue on torch's current stream (graph replay)

Tensors are device tensors of 16-byte residues (torch int64 (n, 2)).  The
plan keeps the pointers, so refill inputs in place between executions.
"""
from __future__ import annotations

import ctypes
import json
import os

from . import _lib
from ._lib import BatchInfo, MulTask, NttTask, TuneEntry, U128
from .ops import _dev_ptr, _nelem, _stream

TUNE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune")


def _log2(n: int) -> int:
    l = n.bit_length() - 1
    if n != 1 << l:
        raise ValueError("transform sizes must be powers of two")
    return l


def load_tune_table(path: str | None = None) -> dict[int, tuple[int, int]]:
    """{log2n: (elog, passes)} from the JSON table written by autotune.py
    (default: tune/gfx950.json next to this file; empty when absent)."""
    path = path or os.path.join(TUNE_DIR, "gfx950.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        d = json.load(f)
    return {int(k): (int(v["elog"]), int(v["passes"])) for k, v in d["entries"].items()}


def _tune_array(tune):
    if not tune:
        return None, 0
    arr = (TuneEntry * len(tune))()
    for i, (l, (elog, passes)) in enumerate(sorted(tune.items())):
        arr[i] = TuneEntry(l, elog, passes, 0)
    return arr, len(tune)


def _task_array(tasks):
    arr = (NttTask * max(1, len(tasks)))()
    for i, (x, y, tw) in enumerate(tasks):
        n = _nelem(tw)
        if _nelem(x) != n or _nelem(y) != n:
            raise ValueError("x, y and tw of a task must have the same length")
        arr[i] = NttTask(_dev_ptr(x), _dev_ptr(y), _dev_ptr(tw), _log2(n), 0)
    return arr


TW_FLAGS = {"global": 0, "lds": 2, "wide": 4}


class BatchPlan:
    """bntt_batch_plan_*: tasks are (x, y, tw) device tensors, muls (a, b, c);
    tw_mode "global" (twiddles from a digit copy in device memory), "lds"
    (staged in LDS where they keep three workgroups per CU) or "wide" (staged
    up to 80 KiB of LDS).  fixed_tables: the twiddle tables are not modified
    after creation, so they are converted once (BNTT_PLAN_TW_FIXED) instead of
    at every execution.  A product whose operands are exactly the outputs of
    two transforms of its size and one table is computed inside their last
    pass."""

    def __init__(self, tasks, muls=(), modulus: int = 0, mu: int = 0, tune=None, graph: bool = True,
                 tw_mode: str = "global", fixed_tables: bool = False):
        if tw_mode not in TW_FLAGS:
            raise ValueError(f"tw_mode must be one of {sorted(TW_FLAGS)}")
        self._keep = (list(tasks), list(muls))  # the buffers must outlive the plan
        tarr = _task_array(self._keep[0])
        marr = (MulTask * max(1, len(self._keep[1])))()
        for i, (a, b, c) in enumerate(self._keep[1]):
            n = _nelem(c)
            if _nelem(a) < n or _nelem(b) < n:
                raise ValueError("product operands shorter than the result")
            marr[i] = MulTask(_dev_ptr(a), _dev_ptr(b), _dev_ptr(c), n)
        tune_arr, ntune = _tune_array(tune)
        handle = ctypes.c_void_p()
        mm, mmu = U128.of(modulus), U128.of(mu)
        _lib.call("bntt_batch_plan_create", ctypes.byref(handle), tarr, len(self._keep[0]), marr,
                  len(self._keep[1]), ctypes.byref(mm), ctypes.byref(mmu), tune_arr, ntune,
                  (_lib.BNTT_PLAN_GRAPH if graph else 0) | TW_FLAGS[tw_mode]
                  | (_lib.BNTT_PLAN_TW_FIXED if fixed_tables else 0))
        self._h = handle

    def execute(self, stream=None) -> None:
        _lib.call("bntt_batch_plan_execute", self._h, _stream(stream))

    def info(self) -> dict:
        inf = BatchInfo()
        _lib.call("bntt_batch_plan_info", self._h, ctypes.byref(inf))
        lv = inf.levels
        return {"levels": lv, "items": inf.items, "tmp_bytes": inf.tmp_bytes,
                "level_items": list(inf.level_items[:lv]), "level_grid": list(inf.level_grid[:lv]),
                "level_lds_bytes": list(inf.level_lds_bytes[:lv])}

    def close(self) -> None:
        if self._h:
            _lib.call("bntt_batch_plan_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TaskList:
    """Transforms as the C array the entry points take (built once)."""

    def __init__(self, tasks):
        self.tasks = list(tasks)
        self.arr = _task_array(self.tasks)

    def __len__(self):
        return len(self.tasks)


def launch_each(tasks, modulus: int, mu: int, tmp=None, stream=None) -> None:
    """One launch sequence per transform (the plain path), from native code.
    `tasks`: a TaskList, or a list of (x, y, tw) to convert first."""
    tl = tasks if isinstance(tasks, TaskList) else TaskList(tasks)
    mm, mmu = U128.of(modulus), U128.of(mu)
    _lib.call("bntt_ntt_launch_each_dev", tl.arr, len(tl), ctypes.byref(mm), ctypes.byref(mmu),
              _dev_ptr(tmp) if tmp is not None else None, _stream(stream))
