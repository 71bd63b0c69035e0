#!/usr/bin/env python3
"""Benchmark of the benchNTT hot path on one MI355X.

Default workload (the BLAS configuration): axpy mod q_B = 2^124 - 59 over
vectors of 2^24 seeded 128-bit residues (256 MiB each), c = a + s*b.  One step
is one axpy call over the whole vectors with inputs resident in HBM.  `value`
is GB/s of algorithmic traffic (48 B per element: two operands in, one out).

The same JSON line also carries vvadd/vvsub/vvmul at the same size, the NTT
rows (plain N=1024 transform per launch, uniform batch 8192 x N=1024 and the
N = 16..512 sweep, the 20,000-transform uneven batch against one launch
sequence per transform), the polynomial-multiply serving trace (saturation
rate, p50/p99 latency over a rate sweep), the roofline of the axpy kernel and
a CPU baseline: the reference's own axpy (oracle/_ref, AVX-512 / AVX2 / scalar
as the host supports) timed on this host's cores over the same vectors.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_ELEM = 48


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_kernel(torch, fn, steps: int, warmup: int, stream) -> float:
    """Average ms per call of fn() over `steps` back-to-back calls, measured
    with HIP events recorded on the stream the kernels are launched on."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def load_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size); see profiles/."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(a_np, b_np, q, mu, s, budget_s: float = 1.5):
    """The reference's axpy on this host's cores over the same 2^24 vectors
    (1024-element calls, as the reference kernel is written): passes for about
    `budget_s` of wall time, which on 16 threads is 10-30 s of CPU work."""
    from oracle import ref

    threads = min(os.cpu_count() or 1, 16)
    if ref.available("ref_blas_scalar.so"):
        isa = ref.best_isa()
        ops = ref.BlasOperands(a_np, b_np, isa)
        n = a_np.shape[0]
        ref.blas("axpy", ops, q, mu, s, threads=threads)  # warm
        t_single = ref.blas("axpy", ops, q, mu, s, threads=1)
        passes, t0 = 0, time.perf_counter()
        tot = 0.0
        while passes < 2 or (time.perf_counter() - t0 < budget_s and passes < 1000):
            tot += ref.blas("axpy", ops, q, mu, s, threads=threads)
            passes += 1
        gbs = n * BYTES_PER_ELEM * passes / tot / 1e9
        return {
            "value": round(gbs, 3),
            "unit": "GB/s",
            "cores": threads,
            "kind": "reference",
            "sample": (f"reference blas_{isa}.c axpy over the 2^24-element vectors in 1024-element calls, "
                       f"{passes} passes, {tot * threads:.0f} s of CPU time, on {threads} threads "
                       f"({ref.cpu_model()}); single-thread "
                       f"{n * BYTES_PER_ELEM / t_single / 1e9:.3f} GB/s"),
        }
    import oracle

    orc = oracle.load()
    m = 1 << 20
    t0 = time.perf_counter()
    orc.axpy(a_np[:m], b_np[:m], s, q, mu)
    dt = time.perf_counter() - t0
    return {"value": round(m * BYTES_PER_ELEM / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "oracle axpy (C restatement) over the first 2^20 elements, 1 thread"}


# VALU instructions per butterfly of the radix-8 register network of the NTT
# kernels (reg_network_fast<3>: 12 butterflies in 1197 VALU instructions on
# the path every wave takes unless a lane's product may reach q, counted in
# the gfx950 ISA of a kernel that loads 8 values and 7 twiddles, runs the
# network and stores; the network's product is 20 v_mad_u64_u32)
VALU_PER_BUTTERFLY = 100


def valu_rates() -> dict:
    """Measured issue rates of the integer VALU instructions (library probe)."""
    import ctypes

    from benchntt_amd import _lib

    out = {}
    for name, op in (("v_mad_u64_u32", 0), ("v_add_co_u32", 1), ("v_cndmask_b32", 2), ("v_lshl_add_u64", 3),
                     ("v_and_b32", 4)):
        r, nom = ctypes.c_double(), ctypes.c_double()
        _lib.call("bntt_probe_valu_rate", op, ctypes.byref(r), ctypes.byref(nom))
        out[name] = r.value
        out["nominal"] = nom.value
    log("VALU issue rates (T lane-ops/s): " + ", ".join(f"{k} {v / 1e12:.2f}" for k, v in out.items()))
    return out


def ntt_roofline(gbfly_per_s: float, rates: dict, what: str) -> dict:
    peak = rates["v_mad_u64_u32"]
    ach = gbfly_per_s * 1e9 * VALU_PER_BUTTERFLY
    return {"bound": "int-valu", "kernel": what, "achieved": round(ach / 1e12, 2), "peak": round(peak / 1e12, 2),
            "unit": "T lane-ops/s", "frac": round(ach / peak, 4), "valu_per_butterfly": VALU_PER_BUTTERFLY,
            "nominal_peak": round(rates["nominal"] / 1e12, 2)}


def cpu_threads() -> int:
    return min(os.cpu_count() or 1, 16)


def ntt_cpu_rows(np, x, tw, qn, mun) -> dict:
    """The reference's ntt() at N = 1024 on this host: one core, 50 timed
    calls after 51 warm-up calls (the reference times 100 calls and averages
    the last 50, ntt_scalar.c:670-683), scalar and the widest AVX build; then
    the uniform batch rate with one transform stream per thread."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ref

    if not ref.available("ref_ntt_scalar_L10.so"):
        return {}
    out = {"ntt_cpu_model": ref.cpu_model()}
    isas = ["scalar"] + ([ref.best_isa()] if ref.best_isa() != "scalar" else [])
    for isa in isas:
        ref.ntt(x, tw, qn, mun, isa=isa, iters=50)
        _, ns = ref.ntt(x, tw, qn, mun, isa=isa, iters=50)
        out[f"ntt_1024_cpu_{isa}_ns_per_butterfly_1core"] = round(ns / 5120, 3)
    isa = isas[-1]
    T = cpu_threads()
    with ThreadPoolExecutor(T) as ex:
        ns = list(ex.map(lambda _: ref.ntt(x, tw, qn, mun, isa=isa, iters=256)[1], range(T)))
    out[f"ntt_uniform_1024_cpu_{isa}_NTTs_per_s"] = round(T / (float(np.mean(ns)) * 1e-9), 1)
    out["ntt_cpu_threads"] = T
    log(f"cpu ntt 1024: " + ", ".join(f"{k.split('_')[3]} {v} ns/butterfly" for k, v in out.items()
                                      if k.endswith("1core")) + f"; {T} threads {isa}: "
        f"{out[f'ntt_uniform_1024_cpu_{isa}_NTTs_per_s']:.0f} NTT/s")
    return out


def uniform_rows(torch, np, bn, qn, mun, stream, per_launch: bool = True, cpu: bool = False) -> dict:
    """Plain transform (N=1024, one launch) and uniform batches: 8192 x N=1024
    and N in {16..512} with 2^23 coefficients, one shared seeded twiddle table
    per N, as one persistent work-queue execution (graph replay) against one
    launch per transform from native code."""
    from benchntt_amd.batch import BatchPlan, TaskList, launch_each, load_tune_table
    from benchntt_amd.residues import uniform

    dev = torch.device("cuda:0")
    out = {}
    tw_np, x_np = uniform(1024, qn, 4), uniform(1024, qn, 3)
    tw = torch.from_numpy(tw_np.view(np.int64)).to(dev)
    x1 = torch.from_numpy(x_np.view(np.int64)).to(dev)
    y1 = torch.empty_like(x1)
    ms1 = time_kernel(torch, lambda: bn.ntt(y1, x1, qn, mun, tw), 200, 20, stream)
    out["ntt_plain_1024_us"] = round(ms1 * 1e3, 2)
    out["ntt_plain_1024_ns_per_butterfly"] = round(ms1 * 1e6 / 5120, 4)
    out["ntt_plain_1024_butterflies_per_s"] = round(5120 / (ms1 * 1e-3), 1)
    # INT-op/s = butterflies/s x the network's VALU lane-instructions per butterfly
    out["ntt_plain_1024_int_ops_per_s"] = round(5120 / (ms1 * 1e-3) * VALU_PER_BUTTERFLY, 1)
    # the reference's call shape: host arrays in, host array out, one call at a
    # time (PCIe and synchronisation included; the library's zero-copy path)
    yh = np.empty_like(x_np)
    for _ in range(20):
        bn.ntt(yh, x_np, qn, mun, tw_np)
    t0 = time.perf_counter()
    for _ in range(200):
        bn.ntt(yh, x_np, qn, mun, tw_np)
    out["ntt_plain_1024_host_entry_us"] = round((time.perf_counter() - t0) / 200 * 1e6, 2)
    log(f"ntt plain 1024: {ms1 * 1e3:.2f} us per launch on device arrays, "
        f"{out['ntt_plain_1024_host_entry_us']} us per host-array call")
    tune = load_tune_table()
    total = 1 << 23
    xb = torch.from_numpy(uniform(total, qn, 5).view(np.int64)).to(dev)
    yb = torch.empty_like(xb)
    s2 = torch.cuda.Stream()
    for L in (10, 4, 5, 6, 7, 8, 9):
        n = 1 << L
        cnt = total // n
        twl = torch.from_numpy(uniform(n, qn, 40 + L).view(np.int64)).to(dev)
        tasks = [(xb[i * n:(i + 1) * n], yb[i * n:(i + 1) * n], twl) for i in range(cnt)]
        with torch.cuda.stream(s2):
            plan = BatchPlan(tasks, [], qn, mun, tune=tune, graph=True, fixed_tables=True)
            msp = time_kernel(torch, plan.execute, 20, 3, s2)
            plan.close()
            if L == 10:
                msd = time_kernel(torch, lambda: bn.ntt(yb, xb, qn, mun, twl), 20, 3, s2)
                tl = TaskList(tasks)
                mse = time_kernel(torch, lambda: launch_each(tl, qn, mun), 2, 1, s2) if per_launch else float("nan")
        key = f"ntt_uniform_{cnt}x{n}"
        out[f"{key}_NTTs_per_s"] = round(cnt / (msp * 1e-3), 1)
        out[f"{key}_Gbutterflies_per_s"] = round(cnt * n * L / 2 / (msp * 1e-3) / 1e9, 2)
        if L == 10:
            out[f"{key}_single_launch_NTTs_per_s"] = round(cnt / (msd * 1e-3), 1)
            out[f"{key}_per_launch_NTTs_per_s"] = round(cnt / (mse * 1e-3), 1)
            out[f"{key}_persistent_vs_per_launch"] = round(mse / msp, 2)
        log(f"uniform {cnt}x{n}: persistent {msp:.4f} ms ({out[f'{key}_NTTs_per_s']:.0f} NTT/s)"
            + (f", one launch {msd:.4f} ms, per-transform launches {mse:.2f} ms" if L == 10 else ""))
    if cpu:
        out.update(ntt_cpu_rows(np, x_np, tw_np, qn, mun))
    return out


def uneven_cpu(np, wl, qn, mun) -> dict:
    """The same 20,000 transforms through the reference's scalar ntt() (one
    build per size), spread over the host's threads: NTTs/s = transforms /
    (sum of the per-transform times, measured concurrently, / threads)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ref

    if not all(ref.available(f"ref_ntt_scalar_L{l}.so") for l in set(int(v) for v in wl.log2n)):
        return {}
    T = cpu_threads()

    def one(i):
        l, off = int(wl.log2n[i]), int(wl.offsets[i])
        return ref.ntt(wl.x[off:off + (1 << l)], wl.tables[l], qn, mun, isa="scalar", iters=1)[1]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(T) as ex:
        ns = float(sum(ex.map(one, range(len(wl.log2n)))))
    wall = time.perf_counter() - t0
    rate = len(wl.log2n) / (ns * 1e-9 / T)
    log(f"cpu uneven 20000 (reference scalar, {T} threads): {rate:.0f} NTT/s, wall {wall:.1f} s")
    return {"uneven_cpu_scalar_NTTs_per_s": round(rate, 1), "uneven_cpu_threads": T}


def uneven_batch(torch, np, bn, qn, mun, stream, per_launch: bool = True, cpu: bool = False) -> dict:
    """20,000 forward NTTs, N log-uniform on 2^4..2^16 in pairs (seeded), one
    twiddle table per size, each pair's pointwise product after its two
    transforms: one persistent work-queue kernel per level, replayed from a
    hipGraph, against one launch sequence per transform from native code."""
    from benchntt_amd.batch import BatchPlan, TaskList, launch_each, load_tune_table
    from benchntt_amd.workloads import uneven_pairs

    wl = uneven_pairs(20000, qn, seed=2024)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(wl.x.view(np.int64)).to(dev)
    dy = torch.empty_like(dx)
    dc = torch.empty((wl.pair_total, 2), dtype=torch.int64, device=dev)
    tabs = {l: torch.from_numpy(t.view(np.int64)).to(dev) for l, t in wl.tables.items()}
    tasks, muls = [], []
    for l, off in zip(wl.log2n, wl.offsets):
        n = 1 << int(l)
        tasks.append((dx[off:off + n], dy[off:off + n], tabs[int(l)]))
    for j, poff in enumerate(wl.pair_offsets):
        n = 1 << int(wl.log2n[2 * j])
        a0, b0 = wl.offsets[2 * j], wl.offsets[2 * j + 1]
        muls.append((dy[a0:a0 + n], dy[b0:b0 + n], dc[poff:poff + n]))
    tune = load_tune_table()
    butterflies = float(sum((1 << int(l)) * int(l) // 2 for l in wl.log2n))
    s2 = torch.cuda.Stream()
    out = {"uneven_total_coefficients": wl.total, "uneven_tuned": bool(tune)}
    with torch.cuda.stream(s2):
        full = BatchPlan(tasks, muls, qn, mun, tune=tune, graph=True, fixed_tables=True)
        only = BatchPlan(tasks, [], qn, mun, tune=tune, graph=True, fixed_tables=True)
        ms_full = time_kernel(torch, full.execute, 5, 2, s2)
        ms_only = time_kernel(torch, only.execute, 5, 2, s2)
        tmp = torch.empty((1 << 16, 2), dtype=torch.int64, device=dev)
        tl = TaskList(tasks)
        ms_each = time_kernel(torch, lambda: launch_each(tl, qn, mun, tmp), 1, 1, s2) if per_launch else float("nan")
    info = full.info()
    full.close()
    only.close()
    out["uneven_batched_NTTs_per_s"] = round(20000 / (ms_only * 1e-3), 1)
    out["uneven_batched_with_products_ms"] = round(ms_full, 3)
    out["uneven_batched_with_products_NTTs_per_s"] = round(20000 / (ms_full * 1e-3), 1)
    out["uneven_per_launch_NTTs_per_s"] = round(20000 / (ms_each * 1e-3), 1)
    out["uneven_batched_vs_per_launch"] = round(ms_each / ms_only, 2)
    out["uneven_Gbutterflies_per_s"] = round(butterflies / (ms_only * 1e-3) / 1e9, 2)
    out["uneven_plan_levels"] = info["levels"]
    log(f"uneven 20000: batched {ms_only:.3f} ms ({out['uneven_batched_NTTs_per_s']:.0f} NTT/s), "
        f"with products {ms_full:.3f} ms, per-launch {ms_each:.1f} ms -> "
        f"{out['uneven_batched_vs_per_launch']}x")
    if cpu:
        del dx, dy, dc, tasks, muls
        out.update(uneven_cpu(np, wl, qn, mun))
    return out


def serving(torch, np, n_req: int = 16384, cpu: bool = True) -> dict:
    """Negacyclic polynomial-multiply requests mod q_116 (forward NTT x2,
    pointwise product, inverse NTT) with N uniform on 2^10..2^16, served with
    continuous batching under a 2 GiB HBM budget on 4 streams.  First a closed
    loop (every request arrived at t = 0) gives the saturation rate; then
    seeded Poisson arrivals at fractions of it give p50/p99 latency (arrival to
    completion seen by the host) and the achieved rate.  16384 requests per
    trace, so that p99 rests on the 164 slowest.  Also: the same
    requests one at a time with no batching.  Operands: seeded uniform
    residues from the library's device generator (bntt_uniform_dev)."""
    from benchntt_amd import config as C
    from benchntt_amd import fixtures as fx
    from benchntt_amd.polymul import PolymulServer

    q = C.Q_116
    rng = np.random.default_rng(55)
    logs = rng.integers(10, 17, size=n_req)
    sizes = [1 << int(l) for l in logs]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    total = int(offs[-1])
    A, B = fx.uniform(total, q, 551), fx.uniform(total, q, 552)
    Cc = torch.empty_like(A)
    reqs = [(A[offs[i]:offs[i + 1]], B[offs[i]:offs[i + 1]], Cc[offs[i]:offs[i + 1]]) for i in range(n_req)]
    torch.cuda.synchronize()
    budget = 2 << 30
    srv = PolymulServer(q, hbm_budget=budget, max_inflight=4, max_batch=64, min_log2n=10, max_log2n=16)
    srv.multiply(reqs[:64])  # warm-up
    out = {"serving_requests": n_req, "serving_hbm_budget_bytes": budget,
           "serving_mean_N": round(float(np.mean(sizes)), 1)}

    def trace(arrival):
        done, st = srv.run_trace(reqs, arrival)
        lat = (done.astype(np.float64) - arrival.astype(np.float64)) * 1e-3  # us
        span = float(done.max()) * 1e-9
        return lat, span, st

    zero = np.zeros(n_req, dtype=np.uint64)
    lat, span, st = trace(zero)
    sat = n_req / span
    out["serving_saturation_requests_per_s"] = round(sat, 1)
    out["serving_saturation_batches"] = int(st["batches"])
    log(f"serving closed loop: {n_req} requests in {span * 1e3:.2f} ms -> {sat:.0f} req/s, "
        f"{st['batches']} batches, at most {st['max_inflight_requests']} requests in flight")
    sweep = []
    for frac in (0.25, 0.5, 0.7, 0.85, 1.0, 1.2):
        rate = frac * sat
        gaps = np.random.default_rng(int(frac * 1000)).exponential(1e9 / rate, size=n_req)
        arrival = np.cumsum(gaps).astype(np.uint64)
        lat, span, st = trace(arrival)
        row = {"offered_per_s": round(rate, 1), "achieved_per_s": round(n_req / span, 1),
               "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
               "batches": int(st["batches"]), "admission_waits": int(st["admission_waits"])}
        sweep.append(row)
        log(f"serving at {frac:.2f} x saturation: offered {rate:.0f}/s achieved {row['achieved_per_s']:.0f}/s, "
            f"p50 {row['p50_us']} us, p99 {row['p99_us']} us, {row['batches']} batches")
    out["serving_sweep"] = sweep
    at = sweep[2]
    out["serving_p50_us_at_0.7_saturation"] = at["p50_us"]
    out["serving_p99_us_at_0.7_saturation"] = at["p99_us"]
    # no batching: one request per bntt_pm_multiply call
    k = 256
    t0 = time.perf_counter()
    for r in reqs[:k]:
        srv.multiply([r])
    one = (time.perf_counter() - t0) / k
    out["serving_unbatched_requests_per_s"] = round(1 / one, 1)
    out["serving_batched_vs_unbatched"] = round(sat * one, 2)
    log(f"serving one request per call: {one * 1e6:.1f} us mean -> {1 / one:.0f} req/s")
    srv.close()
    if cpu:
        import oracle

        orc = oracle.load()
        n = 1 << 12
        from benchntt_amd.residues import uniform

        a, b = uniform(n, q, 1), uniform(n, q, 2)
        t0, cnt = time.perf_counter(), 0
        while time.perf_counter() - t0 < 3.0:
            orc.polymul_ntt(a, b, q)
            cnt += 1
        dt = (time.perf_counter() - t0) / cnt
        out["serving_cpu_baseline"] = {
            "value": round(1 / dt, 2), "unit": "requests/s at N=2^12", "cores": 1, "kind": "port",
            "sample": f"{cnt} negacyclic products of N=2^12 mod q_116 through the exact C restatement "
                      "(bitwise 256-by-128 reduction, table set-up included), about 3 s"}
        log(f"serving cpu port: {1 / dt:.2f} req/s at N=2^12")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-serving", action="store_true", help="leave out the serving trace")
    ap.add_argument("--skip-per-launch", action="store_true",
                    help="skip the one-launch-per-transform comparisons (keeps profiler output small)")
    ap.add_argument("--only", choices=["all", "axpy", "copy", "ntt-uniform", "uneven", "serving"], default="all",
                    help="time one workload only and print no JSON line (for profiling runs): "
                         "axpy 2^24; copy (a 256 MiB device copy, the traffic calibration kernel); "
                         "ntt-uniform (8192 x N=1024); uneven (20,000-transform batch); "
                         "serving (polynomial-multiply trace)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "the transforms and vector ops are independent per element/transform, so a multi-GPU "
              "run is N copies of this one; it is not launched here.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import benchntt_amd as bn
    from benchntt_amd import config as C
    from benchntt_amd.residues import uniform

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    props = torch.cuda.get_device_properties(0)
    log(f"device {props.name} ({getattr(props, 'gcnArchName', '?')}), {props.multi_processor_count} CUs")

    if args.only == "serving":
        serving(torch, np, cpu=False)
        return 0
    if args.only in ("ntt-uniform", "uneven"):
        qn, mun = C.Q_NTT, C.MU_NTT
        if args.only == "uneven":
            uneven_batch(torch, np, bn, qn, mun, stream, not args.skip_per_launch)
            return 0
        uniform_rows(torch, np, bn, qn, mun, stream, not args.skip_per_launch)
        return 0
    if args.only == "copy":
        src = torch.empty(1 << 28, dtype=torch.uint8, device=dev)
        src.fill_(1)
        dst = torch.empty_like(src)
        msc = time_kernel(torch, lambda: dst.copy_(src), args.steps, args.warmup, stream)
        log(f"copy 256 MiB: {msc * 1e3:.1f} us, {2 * (1 << 28) / (msc * 1e-3) / 1e9:.1f} GB/s")
        return 0

    n = 1 << 24
    q, mu, s = C.Q_BLAS, C.MU_BLAS, C.AXPY_SCALAR
    log("generating 3 x 2^24 seeded residues")
    a_np, b_np = uniform(n, q, 1), uniform(n, q, 2)
    da = torch.from_numpy(a_np.view(np.int64)).to(dev)
    db = torch.from_numpy(b_np.view(np.int64)).to(dev)
    dc = torch.empty_like(da)
    torch.cuda.synchronize()

    ms_axpy = time_kernel(torch, lambda: bn.axpy(dc, da, db, s, q, mu), args.steps, args.warmup, stream)
    gbs = n * BYTES_PER_ELEM / (ms_axpy * 1e-3) / 1e9
    log(f"axpy 2^24: {ms_axpy * 1e3:.1f} us/call, {gbs:.1f} GB/s")
    if args.only == "axpy":
        return 0

    extras = {}
    if args.only == "all":
        for name, fn in (("vvadd", lambda: bn.vector_add(dc, da, db, q)),
                         ("vvsub", lambda: bn.vector_sub(dc, da, db, q)),
                         ("vvmul", lambda: bn.vector_mul(dc, da, db, q, mu))):
            ms = time_kernel(torch, fn, args.steps, args.warmup, stream)
            extras[f"{name}_GBps"] = round(n * BYTES_PER_ELEM / (ms * 1e-3) / 1e9, 1)
            log(f"{name} 2^24: {ms * 1e3:.1f} us/call, {extras[f'{name}_GBps']} GB/s")
        # what the device sustains for a plain copy (the runtime's own copy
        # kernel, 256 MiB read + 256 MiB written): the practical HBM ceiling
        src = torch.empty(1 << 28, dtype=torch.uint8, device=dev)
        src.fill_(1)
        dst = torch.empty_like(src)
        msc = time_kernel(torch, lambda: dst.copy_(src), args.steps, args.warmup, stream)
        extras["hbm_device_copy_GBps"] = round(2 * (1 << 28) / (msc * 1e-3) / 1e9, 1)
        extras["axpy_vs_device_copy"] = round(gbs / extras["hbm_device_copy_GBps"], 4)
        log(f"device copy 256 MiB: {extras['hbm_device_copy_GBps']} GB/s")
        del src, dst

        cpu = not args.no_cpu_baseline
        extras.update(uniform_rows(torch, np, bn, C.Q_NTT, C.MU_NTT, stream, not args.skip_per_launch, cpu))
        extras.update(uneven_batch(torch, np, bn, C.Q_NTT, C.MU_NTT, stream, not args.skip_per_launch, cpu))
        if not args.skip_serving:
            extras.update(serving(torch, np, cpu=not args.no_cpu_baseline))
        rates = valu_rates()
        extras["valu_issue_rates_T_lane_ops_per_s"] = {k: round(v / 1e12, 3) for k, v in rates.items()}
        extras["ntt_roofline_uniform_8192x1024"] = ntt_roofline(
            extras["ntt_uniform_8192x1024_single_launch_NTTs_per_s"] * 5120 / 1e9, rates,
            "single-launch kernel, 8192 x N=1024 (ntt_pass_kernel<3>)")
        extras["ntt_roofline_uneven_20000"] = ntt_roofline(
            extras["uneven_Gbutterflies_per_s"], rates, "persistent work-queue kernel, 20,000 transforms 2^4..2^16")

    traffic = load_traffic("axpy_2_24")
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(a_np, b_np, q, mu, s)
        log(f"cpu baseline: {cpu['value']} GB/s on {cpu['cores']} cores ({cpu['kind']})")

    line = {
        "metric": "NTTs/sec batched mixed-N vs per-NTT launch; axpy HBM GB/s; p99 latency",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_axpy, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u128",
        "data": "synthetic (seeded uniform residues in [0, q))",
        "config": {"workload": "axpy c = a + s*b mod (2^124-59), 2^24 residues of 128 bits per vector "
                               "(256 MiB each), inputs resident in HBM"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
