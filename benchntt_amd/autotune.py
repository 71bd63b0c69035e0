"""Autotuner of the batch kernel: per transform size, the radix and pass count.

For each log2n in [lo, hi], every valid (elog, passes) choice -- 2^elog points
per thread per register sub-pass (LDS tile 256 << elog values), and the number
of passes through device memory -- is timed on a uniform batch of that size
(2^22 seeded coefficients, one twiddle table) executed as a batch plan from a
graph; the fastest choice is written with its measured ns per butterfly to
tune/<arch>.json, which BatchPlan users load with load_tune_table().

    python -m benchntt_amd.autotune [--lo 1] [--hi 16] [--out path]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np


def valid(L: int, elog: int, passes: int) -> bool:
    """Same rule as the planner (pass_split in csrc/ntt_core.hpp): the stages
    can be split into `passes` passes of elog..8+elog stages each."""
    tile_log = 8 + elog
    if not (1 <= elog <= 3 and 1 <= passes <= 3):
        return False
    if passes == 1:
        return elog <= L <= tile_log
    lo, hi = elog * passes, tile_log * passes
    return lo <= L <= hi


def candidates(L: int):
    return [(e, p) for e in (1, 2, 3) for p in (1, 2, 3) if valid(L, e, p)]


def time_choice(torch, L: int, elog: int, passes: int, x, y, tw, q: int, mu: int, reps: int) -> float:
    from .batch import BatchPlan

    n = 1 << L
    count = x.shape[0] // n
    tasks = [(x[i * n:(i + 1) * n], y[i * n:(i + 1) * n], tw) for i in range(count)]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        plan = BatchPlan(tasks, [], q, mu, tune={L: (elog, passes)}, graph=True)
        plan.execute()
        plan.execute()
        stream.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(reps):
            plan.execute()
        t1.record(stream)
        stream.synchronize()
    plan.close()
    ms = t0.elapsed_time(t1) / reps
    return ms * 1e6 / (count * n * L / 2)  # ns per butterfly


def tune(lo: int = 1, hi: int = 16, total_log2: int = 22, reps: int = 5, log=print) -> dict:
    import torch

    from . import config as C
    from .residues import uniform

    q, mu = C.Q_NTT, C.MU_NTT
    dev = torch.device("cuda:0")
    total = 1 << total_log2
    x = torch.from_numpy(uniform(total, q, 1).view(np.int64)).to(dev)
    y = torch.empty_like(x)
    entries = {}
    for L in range(lo, hi + 1):
        tw = torch.from_numpy(uniform(1 << L, q, 100 + L).view(np.int64)).to(dev)
        res = {}
        for elog, passes in candidates(L):
            res[f"{elog},{passes}"] = time_choice(torch, L, elog, passes, x, y, tw, q, mu, reps)
        best = min(res, key=res.get)
        elog, passes = (int(v) for v in best.split(","))
        entries[str(L)] = {"elog": elog, "passes": passes, "ns_per_butterfly": round(res[best], 5),
                           "candidates": {k: round(v, 5) for k, v in res.items()}}
        log(f"log2n {L:2d}: best elog {elog} passes {passes} at {res[best]:.4f} ns/butterfly "
            f"({len(res)} candidates)")
    props = torch.cuda.get_device_properties(0)
    return {"device": props.name, "arch": getattr(props, "gcnArchName", ""), "total_coefficients": total,
            "entries": entries}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lo", type=int, default=1)
    ap.add_argument("--hi", type=int, default=16)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune",
                                                  "gfx950.json"))
    args = ap.parse_args()
    table = tune(args.lo, args.hi, reps=args.reps, log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
