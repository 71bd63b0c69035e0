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
    """The reference's axpy on every core this process may use, over the same
    2^24 vectors (1024-element calls, as the reference kernel is written; the
    reference's own build for the host's ISA, AVX-512 / AVX2 / scalar):
    passes for about `budget_s` of wall time, 10-30 s of CPU work on 16 cores."""
    from oracle import ref

    threads = ref.cpu_threads()
    if ref.available("ref_blas_scalar.so"):
        isa, fl = ref.best_isa(), ref.flavour()
        ops = ref.BlasOperands(a_np, b_np, isa, fl=fl)
        n = a_np.shape[0]
        ref.blas("axpy", ops, q, mu, s, threads=threads)  # warm
        t_single = ref.blas("axpy", ops, q, mu, s, threads=1)
        passes, t0 = 0, time.perf_counter()
        tot = 0.0
        while passes < 2 or (time.perf_counter() - t0 < budget_s and passes < 1000):
            tot += ref.blas("axpy", ops, q, mu, s, threads=threads)
            passes += 1
        gbs = n * BYTES_PER_ELEM * passes / tot / 1e9
        host = ref.host_info()
        return {
            "value": round(gbs, 3),
            "unit": "GB/s",
            "cores": threads,
            "kind": "reference",
            "sample": (f"reference blas_{isa}.c axpy over the 2^24-element vectors in 1024-element calls, "
                       f"{passes} passes, {tot * threads:.0f} s of CPU time, on {threads} threads (every core "
                       f"this process may use: nproc {host['nproc']}, affinity {host['affinity']}); "
                       f"single-thread {n * BYTES_PER_ELEM / t_single / 1e9:.3f} GB/s"),
            "host": host,
        }
    import oracle

    orc = oracle.load()
    m = 1 << 20
    t0 = time.perf_counter()
    orc.axpy(a_np[:m], b_np[:m], s, q, mu)
    dt = time.perf_counter() - t0
    return {"value": round(m * BYTES_PER_ELEM / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "oracle axpy (C restatement) over the first 2^20 elements, 1 thread"}


def load_valu_counts() -> dict:
    """VALU instructions per butterfly of the NTT kernels, per workload, from
    rocprofv3 SQ counters (SQ_INSTS_VALU x 64 lanes / butterflies; see
    profiles/ntt_valu_counts.json and the command that produced it)."""
    path = os.path.join(ROOT, "profiles", "ntt_valu_counts.json")
    try:
        with open(path) as f:
            return json.load(f).get("workloads", {})
    except (OSError, ValueError):
        return {}


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


def ntt_roofline(gbfly_per_s: float, rates: dict, what: str, key: str) -> dict:
    """INT-VALU roofline of an NTT workload: achieved lane-ops/s = butterflies/s
    x the VALU lane-instructions per butterfly counted for that workload
    (profiles/ntt_valu_counts.json), against the nominal issue rate of one
    VALU lane-instruction per lane per clock (CUs x 64 x clock); the measured
    v_mad_u64_u32 issue rate, the product's own instruction, is given beside."""
    counts = load_valu_counts().get(key)
    nominal = rates["nominal"]
    out = {"bound": "int-valu", "kernel": what, "unit": "T lane-ops/s", "peak": round(nominal / 1e12, 2),
           "peak_v_mad_u64_u32_measured": round(rates["v_mad_u64_u32"] / 1e12, 2)}
    if not counts:
        out.update({"achieved": None, "frac": None, "valu_per_butterfly": None,
                    "note": f"no SQ counts for {key} in profiles/ntt_valu_counts.json"})
        return out
    vpb = counts["valu_per_butterfly"]
    ach = gbfly_per_s * 1e9 * vpb
    out.update({"achieved": round(ach / 1e12, 2), "frac": round(ach / nominal, 4),
                "frac_of_v_mad_rate": round(ach / rates["v_mad_u64_u32"], 4), "valu_per_butterfly": vpb,
                "valu_per_butterfly_source": counts.get("source", "profiles/ntt_valu_counts.json")})
    return out


def cpu_threads() -> int:
    from oracle import ref

    return ref.cpu_threads()


def ntt_sweep(torch, np, qn, mun, cpu: bool = True) -> dict:
    """The reference's own NTT measurement (benchmark/run_ntt.sh:13, N = 2^8 ..
    2^20; ntt_scalar.c:670-683,719: ns per butterfly of one ntt() call, mean of
    the last 50 of 100 calls): the plain one-transform GPU path on device
    arrays (bntt_ntt_fwd_dev, count 1, back-to-back calls timed with HIP
    events, and the same calls replayed from a captured graph, which leaves
    out the host's cost of each call), the drop-in on host arrays in the
    reference's protocol (bntt_ntt_fwd timed per call in C, PCIe and
    synchronisation included), and the reference's scalar, AVX2 and AVX-512
    ntt() on one host core (its AMD build where the host has the ISA).  Large
    N take fewer calls (still the mean of the last half), so the sweep stays
    within a few seconds.  Seeded uniform inputs and a seeded table per N."""
    import ctypes

    from benchntt_amd import _lib, calltime
    from benchntt_amd._lib import U128
    from benchntt_amd.residues import uniform

    lib = _lib.load()
    mq, mm = U128.of(qn), U128.of(mun)
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream()
    rows = []
    have_ref = False
    if cpu:
        from oracle import ref

        have_ref = ref.available("ref_ntt_scalar_L8.so")
        isas = ["scalar", "avx2"] + (["avx512"] if ref.best_isa() == "avx512" else [])
        fl = ref.flavour()
    for L in range(8, 21):
        n = 1 << L
        bfly = n * L // 2
        x_np, tw_np = uniform(n, qn, 3 + L), uniform(n, qn, 40 + L)
        x = torch.from_numpy(x_np.view(np.int64)).to(dev)
        tw = torch.from_numpy(tw_np.view(np.int64)).to(dev)
        y = torch.empty_like(x)
        need = lib.bntt_ntt_tmp_bytes(L, 1) // 16
        tmp = torch.empty((max(need, 1), 2), dtype=torch.int64, device=dev)
        args = (y.data_ptr(), x.data_ptr(), tw.data_ptr(), L, 1, ctypes.byref(mq), ctypes.byref(mm),
                tmp.data_ptr() if need else None, ctypes.c_void_p(s.cuda_stream))
        calls = 100 if L <= 16 else 20
        for _ in range(calls // 2):
            _lib.check(lib.bntt_ntt_fwd_dev(*args))
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(s)
        for _ in range(calls // 2):
            _lib.check(lib.bntt_ntt_fwd_dev(*args))
        ev1.record(s)
        torch.cuda.synchronize()
        us = ev0.elapsed_time(ev1) / (calls // 2) * 1e3
        # the same calls replayed from a captured graph: the device's time per
        # transform without the host's cost of issuing each call
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(20):
                _lib.check(lib.bntt_ntt_fwd_dev(*args[:-1], ctypes.c_void_p(s.cuda_stream)))
        g.replay()
        torch.cuda.synchronize()
        ev0.record(s)
        with torch.cuda.stream(s):
            for _ in range(5):
                g.replay()
        ev1.record(s)
        torch.cuda.synchronize()
        gus = ev0.elapsed_time(ev1) / 100 * 1e3
        del g
        # the reference's call shape and protocol: host arrays in and out, each
        # call timed in C, mean of the last half; its output must equal the
        # device path's
        yh = np.empty_like(x_np)
        hns = calltime.ntt_ns_per_call(yh, x_np, tw_np, qn, mun, iters=calls)
        if not np.array_equal(yh, y.cpu().numpy().view(np.uint64).reshape(yh.shape)):
            raise RuntimeError(f"host-array ntt() at N=2^{L} differs from the device path")
        row = {"log2n": L, "gpu_plain_us": round(us, 2), "gpu_plain_ns_per_butterfly": round(us * 1e3 / bfly, 4),
               "gpu_plain_graph_us": round(gus, 2), "gpu_plain_graph_ns_per_butterfly": round(gus * 1e3 / bfly, 4),
               "gpu_host_entry_us": round(hns / 1e3, 2), "gpu_host_entry_ns_per_butterfly": round(hns / bfly, 4),
               "gpu_host_entry_calls": calls}
        if have_ref:
            for isa in isas:
                _, ns1 = ref.ntt(x_np, tw_np, qn, mun, isa=isa, iters=2, fl=fl)
                iters = int(min(100, max(4, 0.15e9 / max(ns1, 1.0))))
                iters -= iters % 2
                _, ns = ref.ntt(x_np, tw_np, qn, mun, isa=isa, iters=iters, fl=fl)
                row[f"cpu_{isa}_ns_per_butterfly"] = round(ns / bfly, 4)
                row[f"cpu_{isa}_calls"] = iters
        rows.append(row)
        log(f"sweep N=2^{L}: GPU plain {us:.2f} us ({row['gpu_plain_ns_per_butterfly']} ns/butterfly), "
            f"{gus:.2f} us from a graph, host arrays {hns / 1e3:.2f} us ({row['gpu_host_entry_ns_per_butterfly']} "
            "ns/butterfly)"
            + "".join(f", {k.split('_')[1]} {v} ns/butterfly" for k, v in row.items()
                      if k.startswith("cpu_") and k.endswith("butterfly")))
    out = {"ntt_sweep": rows}
    if have_ref:
        out["ntt_sweep_cpu"] = {"cores": 1, "flavour": fl, "compiler_flags": ref.FLAVOUR_FLAGS[fl],
                                "model": ref.host_info()["model"]}
    return out


def blas_1024_rows(torch, np, bn, q, mu, s, cpu: bool = True) -> dict:
    """The reference's BLAS measurement at its own size (blas_scalar.c:660-721,
    777-779: ns per element of one 1024-element call, mean of the last 50 of
    100 calls; the AVX-512 add and sub take the last 500 of 1000): each op on
    the GPU through device arrays (back-to-back launches, HIP events) and
    through the reference's call shape (host arrays in and out, one call at a
    time, timed per call in C as the reference does, and from Python), beside
    the reference's scalar, AVX2 and AVX-512 kernels on one host core.
    Operands: seeded uniform residues mod q_B."""
    from benchntt_amd import calltime
    from benchntt_amd.residues import uniform

    n = 1024
    a_np, b_np = uniform(n, q, 11), uniform(n, q, 12)
    dev = torch.device("cuda:0")
    da = torch.from_numpy(a_np.view(np.int64)).to(dev)
    db = torch.from_numpy(b_np.view(np.int64)).to(dev)
    dc = torch.empty_like(da)
    ch = np.empty_like(a_np)
    st = torch.cuda.current_stream()
    fns = {
        "vvadd": (lambda: bn.vector_add(dc, da, db, q), lambda: bn.vector_add(ch, a_np, b_np, q)),
        "vvsub": (lambda: bn.vector_sub(dc, da, db, q), lambda: bn.vector_sub(ch, a_np, b_np, q)),
        "vvmul": (lambda: bn.vector_mul(dc, da, db, q, mu), lambda: bn.vector_mul(ch, a_np, b_np, q, mu)),
        "axpy": (lambda: bn.axpy(dc, da, db, s, q, mu), lambda: bn.axpy(ch, a_np, b_np, s, q, mu)),
    }
    out = {}
    ref_ops = {}
    if cpu:
        from oracle import ref

        if ref.available("ref_blas_scalar.so"):
            fl = ref.flavour()
            isas = ["scalar", "avx2"] + (["avx512"] if ref.best_isa() == "avx512" else [])
            ref_ops = {isa: ref.BlasOperands(a_np, b_np, isa, fl=fl) for isa in isas}
    for op, (fdev, fhost) in fns.items():
        ms = time_kernel(torch, fdev, 50, 50, st)
        row = {"gpu_device_ns_per_element": round(ms * 1e6 / n, 3)}
        # host arrays, timed per call in C in the reference's protocol; the
        # output must equal the device path's
        ns = calltime.blas_ns_per_call(op, ch, a_np, b_np, q, mu, s, iters=100)
        row["gpu_host_entry_ns_per_element"] = round(ns / n, 3)
        fdev()
        torch.cuda.synchronize()
        if not np.array_equal(ch, dc.cpu().numpy().view(np.uint64).reshape(ch.shape)):
            raise RuntimeError(f"host-array {op} differs from the device path")
        for _ in range(20):
            fhost()
        t0 = time.perf_counter()
        for _ in range(100):
            fhost()
        row["gpu_host_entry_python_ns_per_element"] = round((time.perf_counter() - t0) / 100 * 1e9 / n, 3)
        for isa, ops in ref_ops.items():
            iters = 1000 if (isa == "avx512" and op in ("vvadd", "vvsub")) else 100
            row[f"cpu_{isa}_ns_per_element"] = round(ref.blas_ns_per_element(op, ops, q, mu, s, iters), 3)
        out[op] = row
        log(f"blas {op} n=1024: " + ", ".join(f"{k} {v}" for k, v in row.items()))
    return {"blas_1024": out}


def uniform_rows(torch, np, bn, qn, mun, stream, per_launch: bool = True, cpu: bool = False) -> dict:
    """Plain transform (N=1024, one launch) and uniform batches: 8192 x N=1024
    and N in {16..512} with 2^23 coefficients, one shared seeded twiddle table
    per N, as one persistent work-queue execution (graph replay) against one
    launch per transform from native code."""
    import ctypes

    from benchntt_amd import _lib, calltime
    from benchntt_amd.batch import BatchPlan, TaskList, launch_each, load_tune_table
    from benchntt_amd.residues import uniform

    dev = torch.device("cuda:0")
    out = {}
    tw_np, x_np = uniform(1024, qn, 4), uniform(1024, qn, 3)
    tw = torch.from_numpy(tw_np.view(np.int64)).to(dev)
    x1 = torch.from_numpy(x_np.view(np.int64)).to(dev)
    y1 = torch.empty_like(x1)
    # back-to-back calls of the C entry point on device arrays (HIP events on
    # the launch stream), and the same calls through the Python wrapper, whose
    # own cost per call comes close to the kernel's
    lib = _lib.load()
    mq, mm = _lib.U128.of(qn), _lib.U128.of(mun)
    args = (y1.data_ptr(), x1.data_ptr(), tw.data_ptr(), 10, 1, ctypes.byref(mq), ctypes.byref(mm), None,
            ctypes.c_void_p(stream.cuda_stream))
    fwd = lib.bntt_ntt_fwd_dev
    ms1 = time_kernel(torch, lambda: fwd(*args), 200, 20, stream)
    _lib.check(fwd(*args))
    out["ntt_plain_1024_us"] = round(ms1 * 1e3, 2)
    out["ntt_plain_1024_ns_per_butterfly"] = round(ms1 * 1e6 / 5120, 4)
    out["ntt_plain_1024_butterflies_per_s"] = round(5120 / (ms1 * 1e-3), 1)
    msp = time_kernel(torch, lambda: bn.ntt(y1, x1, qn, mun, tw), 200, 20, stream)
    out["ntt_plain_1024_python_us"] = round(msp * 1e3, 2)
    # the reference's call shape: host arrays in, host array out, one call at a
    # time (PCIe and synchronisation included; the library's zero-copy path),
    # timed per call in C as the reference times ntt() (ntt_scalar.c:671-684:
    # mean of the last 50 of 100 calls), and the same calls from Python
    # (bn.ntt on numpy arrays)
    yh = np.empty_like(x_np)
    out["ntt_plain_1024_host_entry_us"] = round(calltime.ntt_ns_per_call(yh, x_np, tw_np, qn, mun, 100) / 1e3, 2)
    if not np.array_equal(yh, y1.cpu().numpy().view(np.uint64).reshape(yh.shape)):
        raise RuntimeError("host-array ntt() at N=1024 differs from the device path")
    for _ in range(20):
        bn.ntt(yh, x_np, qn, mun, tw_np)
    t0 = time.perf_counter()
    for _ in range(200):
        bn.ntt(yh, x_np, qn, mun, tw_np)
    out["ntt_plain_1024_host_entry_python_us"] = round((time.perf_counter() - t0) / 200 * 1e6, 2)
    log(f"ntt plain 1024: {ms1 * 1e3:.2f} us per launch on device arrays ({msp * 1e3:.2f} from Python), "
        f"{out['ntt_plain_1024_host_entry_us']} us per host-array call timed in C, "
        f"{out['ntt_plain_1024_host_entry_python_us']} us from Python")
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
            msp = time_kernel(torch, plan.execute, 200, 200, s2)
            plan.close()
            if L == 10:
                msd = time_kernel(torch, lambda: bn.ntt(yb, xb, qn, mun, twl), 200, 200, s2)
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
        out.update(uniform_cpu(np, x_np, tw_np, qn, mun))
    return out


def uniform_cpu(np, x, tw, qn, mun) -> dict:
    """The uniform batch on the CPU: the reference's fastest ntt() build for
    the host at N = 1024, one transform stream per core, every core this
    process may use."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ref

    if not ref.available("ref_ntt_scalar_L10.so"):
        return {}
    isa, fl = ref.best_isa(), ref.flavour()
    T = cpu_threads()
    ref.ntt(x, tw, qn, mun, isa=isa, iters=4, fl=fl)
    with ThreadPoolExecutor(T) as ex:
        ns = list(ex.map(lambda _: ref.ntt(x, tw, qn, mun, isa=isa, iters=256, fl=fl)[1], range(T)))
    rate = T / (float(np.mean(ns)) * 1e-9)
    log(f"cpu uniform 1024 ({isa}, {T} threads): {rate:.0f} NTT/s")
    return {f"ntt_uniform_1024_cpu_{isa}_NTTs_per_s": round(rate, 1), "ntt_cpu_threads": T}


def uneven_cpu(np, wl, qn, mun) -> dict:
    """The same 20,000 transforms through the reference's fastest ntt() build
    for the host (AVX-512 where available, one build per size), spread over
    every core this process may use: NTTs/s = transforms / (sum of the
    per-transform times, measured concurrently, / threads)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ref

    isa, fl = ref.best_isa(), ref.flavour()
    sfx = "_amd" if fl == "amd" else ""
    if not all(ref.available(f"ref_ntt_{isa}{sfx}_L{l}.so") for l in set(int(v) for v in wl.log2n)):
        return {}
    T = cpu_threads()

    def one(i):
        l, off = int(wl.log2n[i]), int(wl.offsets[i])
        return ref.ntt(wl.x[off:off + (1 << l)], wl.tables[l], qn, mun, isa=isa, iters=2, fl=fl)[1]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(T) as ex:
        ns = float(sum(ex.map(one, range(len(wl.log2n)))))
    wall = time.perf_counter() - t0
    rate = len(wl.log2n) / (ns * 1e-9 / T)
    log(f"cpu uneven 20000 (reference {isa}, {T} threads): {rate:.0f} NTT/s, wall {wall:.1f} s")
    return {f"uneven_cpu_{isa}_NTTs_per_s": round(rate, 1), "uneven_cpu_threads": T}


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
        ms_full = time_kernel(torch, full.execute, 20, 10, s2)
        ms_only = time_kernel(torch, only.execute, 20, 10, s2)
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


def serving(torch, np, n_req: int = 16384) -> dict:
    """Negacyclic polynomial-multiply requests mod q_116 (forward NTT x2,
    pointwise product, inverse NTT) with N uniform on 2^10..2^16, served with
    continuous batching under a 2 GiB HBM budget on 3 streams, batches closed
    at 64 requests or 20 M butterflies.  First a closed
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
    streams = 3
    # batches close at 64 requests or 20 M butterflies of work, whichever
    # comes first (a batch's duration follows its work)
    cap = 20_000_000
    srv = PolymulServer(q, hbm_budget=budget, max_inflight=streams, max_batch=64, min_log2n=10, max_log2n=16,
                        max_batch_butterflies=cap)
    srv.multiply(reqs[:64])  # warm-up
    out = {"serving_requests": n_req, "serving_hbm_budget_bytes": budget, "serving_max_batch": 64,
           "serving_max_batch_butterflies": cap, "serving_streams": streams,
           "serving_mean_N": round(float(np.mean(sizes)), 1)}

    def trace(arrival):
        done, st = srv.run_trace(reqs, arrival)
        lat = (done.astype(np.float64) - arrival.astype(np.float64)) * 1e-3  # us
        span = float(done.max()) * 1e-9
        return lat, span, st

    zero = np.zeros(n_req, dtype=np.uint64)
    # untimed warm-up: one closed loop and one Poisson trace
    trace(zero)
    trace(np.cumsum(np.random.default_rng(9).exponential(1e9 / 3e5, size=n_req)).astype(np.uint64))
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
               "batches": int(st["batches"]), "admission_waits": int(st["admission_waits"]),
               "max_batch_us": round(st["max_batch_ns"] * 1e-3, 1),
               "max_host_loop_gap_us": round(st["max_loop_gap_ns"] * 1e-3, 1)}
        sweep.append(row)
        log(f"serving at {frac:.2f} x saturation: offered {rate:.0f}/s achieved {row['achieved_per_s']:.0f}/s, "
            f"p50 {row['p50_us']} us, p99 {row['p99_us']} us, {row['batches']} batches, "
            f"longest batch {row['max_batch_us']} us, longest host loop gap {row['max_host_loop_gap_us']} us")
    out["serving_sweep"] = sweep
    at = sweep[2]
    out["serving_p50_us_at_0.7_saturation"] = at["p50_us"]
    out["serving_p99_us_at_0.7_saturation"] = at["p99_us"]
    # admission under a budget that binds: the 0.7 x saturation trace again
    # with a 128 MiB arena (about 100 requests of the mean size), results
    # compared on the device with the unconstrained run's (outside the timing)
    ref_c = Cc.clone()
    tight = 128 << 20
    srv_t = PolymulServer(q, hbm_budget=tight, max_inflight=streams, max_batch=64, min_log2n=10, max_log2n=16,
                          max_batch_butterflies=cap)
    srv_t.multiply(reqs[:4])
    Cc.zero_()
    gaps = np.random.default_rng(700).exponential(1e9 / (0.7 * sat), size=n_req)
    arrival = np.cumsum(gaps).astype(np.uint64)
    done, st = srv_t.run_trace(reqs, arrival)
    lat = (done.astype(np.float64) - arrival.astype(np.float64)) * 1e-3
    span = float(done.max()) * 1e-9
    same = bool(torch.equal(Cc, ref_c))
    out["serving_budget_bound"] = {
        "hbm_budget_bytes": tight, "offered_per_s": round(0.7 * sat, 1), "achieved_per_s": round(n_req / span, 1),
        "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
        "admission_waits": int(st["admission_waits"]), "max_inflight_requests": int(st["max_inflight_requests"]),
        "batches": int(st["batches"]), "results_equal_unconstrained": same}
    log(f"serving with a {tight >> 20} MiB budget at 0.7 x saturation: achieved {n_req / span:.0f}/s, "
        f"p50 {out['serving_budget_bound']['p50_us']} us, p99 {out['serving_budget_bound']['p99_us']} us, "
        f"{st['admission_waits']} admission waits, results equal: {same}")
    srv_t.close()
    del ref_c
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
    ap.add_argument("--only", choices=["all", "axpy", "vvadd", "vvsub", "vvmul", "copy", "ntt-uniform", "uneven",
                                       "serving", "sweep", "blas1024"],
                    default="all",
                    help="time one workload only and print no JSON line (for profiling runs): "
                         "axpy, vvadd, vvsub or vvmul at 2^24; copy (a 256 MiB device copy, the traffic calibration kernel); "
                         "ntt-uniform (8192 x N=1024); uneven (20,000-transform batch); "
                         "serving (polynomial-multiply trace); sweep (plain NTT N = 2^8..2^20 beside the "
                         "reference's CPU ntt()); blas1024 (the reference's n = 1024 BLAS protocol)")
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
        serving(torch, np)
        return 0
    if args.only == "sweep":
        ntt_sweep(torch, np, C.Q_NTT, C.MU_NTT, cpu=not args.no_cpu_baseline)
        return 0
    if args.only == "blas1024":
        blas_1024_rows(torch, np, bn, C.Q_BLAS, C.MU_BLAS, C.AXPY_SCALAR, cpu=not args.no_cpu_baseline)
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

    vv = {"vvadd": lambda: bn.vector_add(dc, da, db, q), "vvsub": lambda: bn.vector_sub(dc, da, db, q),
          "vvmul": lambda: bn.vector_mul(dc, da, db, q, mu)}
    if args.only in vv:
        ms = time_kernel(torch, vv[args.only], args.steps, args.warmup, stream)
        log(f"{args.only} 2^24: {ms * 1e3:.1f} us/call, {n * BYTES_PER_ELEM / (ms * 1e-3) / 1e9:.1f} GB/s")
        return 0
    ms_axpy = time_kernel(torch, lambda: bn.axpy(dc, da, db, s, q, mu), args.steps, args.warmup, stream)
    gbs = n * BYTES_PER_ELEM / (ms_axpy * 1e-3) / 1e9
    log(f"axpy 2^24: {ms_axpy * 1e3:.1f} us/call, {gbs:.1f} GB/s")
    if args.only == "axpy":
        return 0

    extras = {}
    if args.only == "all":
        for name, fn in vv.items():
            ms = time_kernel(torch, fn, args.steps, args.warmup, stream)
            extras[f"{name}_GBps"] = round(n * BYTES_PER_ELEM / (ms * 1e-3) / 1e9, 1)
            extras[f"{name}_hbm_bytes_per_launch"] = load_traffic(f"{name}_2_24")
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
        extras.update(blas_1024_rows(torch, np, bn, q, mu, s, cpu))
        extras.update(ntt_sweep(torch, np, C.Q_NTT, C.MU_NTT, cpu))
        extras.update(uniform_rows(torch, np, bn, C.Q_NTT, C.MU_NTT, stream, not args.skip_per_launch, cpu))
        extras.update(uneven_batch(torch, np, bn, C.Q_NTT, C.MU_NTT, stream, not args.skip_per_launch, cpu))
        if not args.skip_serving:
            extras.update(serving(torch, np))
        rates = valu_rates()
        extras["valu_issue_rates_T_lane_ops_per_s"] = {k: round(v / 1e12, 3) for k, v in rates.items()}
        extras["ntt_roofline_uniform_8192x1024"] = ntt_roofline(
            extras["ntt_uniform_8192x1024_NTTs_per_s"] * 5120 / 1e9, rates,
            "batch plan (queue_kernel<QK_FAST8>), 8192 x N=1024, whole execution", "uniform_8192x1024_plan")
        extras["ntt_roofline_uniform_8192x1024_single_launch"] = ntt_roofline(
            extras["ntt_uniform_8192x1024_single_launch_NTTs_per_s"] * 5120 / 1e9, rates,
            "single-launch kernel (ntt_pass_kernel<3>), 8192 x N=1024", "uniform_8192x1024_single_launch")
        extras["ntt_roofline_uneven_20000"] = ntt_roofline(
            extras["uneven_Gbutterflies_per_s"], rates,
            "batch plan (queue_kernel<QK_FAST8>), 20,000 transforms 2^4..2^16, whole execution", "uneven_20000")

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
