"""GPU parity of the polynomial-multiply server (negacyclic c = a * b mod
(x^N + 1, q)): products against the oracle (schoolbook up to N = 2^9, the
oracle's exact NTT pipeline above that), closed forms, one request alone
against the same request inside a continuous-batching trace, admission under a
tight HBM budget, and argument errors."""
import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints, to_ints, uniform
from conftest import from_dev, to_dev
from test_oracle import POLY_MODULI

pytestmark = pytest.mark.gpu

UNIT = 256


def req_bytes(L):
    """Arena bytes one request of 2^L points holds while in flight: two
    buffers of N residues below 2^10 points (always one-pass transforms), four
    from 2^10 on (room for two passes)."""
    bufs = 2 if L <= 9 else 4
    return -(-(bufs << L) * 16 // UNIT) * UNIT


def rand(n, q, seed):
    if q.bit_length() > 64:
        return uniform(n, q, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros((n, 2), dtype=np.uint64)
    out[:, 0] = rng.integers(0, q, size=n, dtype=np.uint64)
    return out


def server(q, **kw):
    from benchntt_amd.polymul import PolymulServer

    print(f"\nserver q={q:#x} {kw}", flush=True)
    return PolymulServer(q, **kw)


def run_batch(torch, srv, pairs):
    """One bntt_pm_multiply batch; returns host results."""
    dev = [(to_dev(torch, a), to_dev(torch, b)) for a, b in pairs]
    outs = [torch.empty_like(da) for da, _ in dev]
    srv.multiply([(da, db, dc) for (da, db), dc in zip(dev, outs)])
    return [from_dev(c) for c in outs]


def expect(orc, a, b, q):
    n = a.shape[0]
    return orc.polymul_schoolbook(a, b, q, threads=16) if n <= 4096 else orc.polymul_ntt(a, b, q)


@pytest.mark.parametrize("name", sorted(POLY_MODULI))
def test_small_sizes_schoolbook(gpu, orc, name):
    """Every size 2^1 .. 2^9 (at most the modulus' 2-adicity), two random
    requests each plus edge operands, all in one batch, against schoolbook."""
    q, lmax = POLY_MODULI[name]
    top = min(lmax, 9)
    srv = server(q, min_log2n=1, max_log2n=top, max_batch=64)
    pairs = []
    for L in range(1, top + 1):
        n = 1 << L
        pairs.append((rand(n, q, 10 * L), rand(n, q, 10 * L + 1)))
        pairs.append((rand(n, q, 10 * L + 2), rand(n, q, 10 * L + 3)))
        pairs.append((from_ints([q - 1] * n), from_ints([q - 1] * n)))
    got = run_batch(gpu, srv, pairs)
    for i, ((a, b), c) in enumerate(zip(pairs, got)):
        assert np.array_equal(c, orc.polymul_schoolbook(a, b, q)), f"request {i} (N={a.shape[0]})"
    srv.close()


def test_schoolbook_2_10_to_2_12(gpu, orc):
    """The serving sizes up to 2^12 (single- and two-pass transforms) over
    q_116 against schoolbook convolution."""
    q = C.Q_116
    srv = server(q, min_log2n=10, max_log2n=12, max_batch=8)
    pairs = [(rand(1 << L, q, 300 + L), rand(1 << L, q, 400 + L)) for L in (10, 11, 12)]
    got = run_batch(gpu, srv, pairs)
    for (a, b), c in zip(pairs, got):
        assert np.array_equal(c, orc.polymul_schoolbook(a, b, q, threads=16)), f"N={a.shape[0]}"
    srv.close()


@pytest.mark.parametrize("name", ["q116", "q122", "q64"])
def test_closed_forms(gpu, name):
    """x * x^(N-1) = -1, 1 * b = b, 0 * b = 0, x^i * x^j = +-x^((i+j) mod N)
    (pins the twist, bit-reversal and sign of the negacyclic wrap)."""
    q, _ = POLY_MODULI[name]
    srv = server(q, min_log2n=3, max_log2n=16, max_batch=64)
    rng = np.random.default_rng(5)
    pairs, want = [], []
    for L in (3, 8, 11, 12, 13, 16):
        n = 1 << L
        b = rand(n, q, L)
        one = from_ints([1] + [0] * (n - 1))
        pairs += [(one, b), (from_ints([0] * n), b)]
        want += [b, from_ints([0] * n)]
        for _ in range(2):
            i, j = (int(v) for v in rng.integers(0, n, size=2))
            xi, xj = from_ints([0] * n), from_ints([0] * n)
            xi[i, 0] = xj[j, 0] = 1
            w = [0] * n
            w[(i + j) % n] = 1 if i + j < n else q - 1
            pairs.append((xi, xj))
            want.append(from_ints(w))
    got = run_batch(gpu, srv, pairs)
    for k, (c, w) in enumerate(zip(got, want)):
        assert np.array_equal(c, w), f"case {k} (N={w.shape[0]})"
    srv.close()


@pytest.mark.parametrize("name,sizes", [("q116", range(10, 17)), ("q122", (12, 16)), ("q64", (11, 16))])
def test_large_sizes_oracle(gpu, orc, name, sizes):
    """Sizes 2^10 .. 2^16 (single- and two-pass transforms) against the
    oracle's exact NTT pipeline, all in one batch."""
    q, _ = POLY_MODULI[name]
    srv = server(q, min_log2n=10, max_log2n=16, max_batch=16)
    pairs = [(rand(1 << L, q, 100 + L), rand(1 << L, q, 200 + L)) for L in sizes]
    got = run_batch(gpu, srv, pairs)
    for (a, b), c in zip(pairs, got):
        assert np.array_equal(c, orc.polymul_ntt(a, b, q)), f"N={a.shape[0]}"
    srv.close()


@pytest.mark.parametrize("name", ["q116", "q122", "q64"])
def test_small_sizes_one_and_two_passes(gpu, orc, name):
    """Sizes 2^10 and 2^11 run in two passes in a small batch and in one pass
    in a batch large enough for 2048-value items (here beside fifteen 2^16
    requests); both equal the oracle's exact NTT pipeline, for the sparse
    (q_116) and dense Montgomery reductions."""
    q, _ = POLY_MODULI[name]
    srv = server(q, min_log2n=10, max_log2n=16, max_batch=16)
    big = [(rand(1 << 16, q, 900 + i), rand(1 << 16, q, 950 + i)) for i in range(15)]
    for L in (10, 11):
        a, b = rand(1 << L, q, 500 + L), rand(1 << L, q, 600 + L)
        want = orc.polymul_ntt(a, b, q)
        assert np.array_equal(run_batch(gpu, srv, [(a, b)])[0], want), f"N=2^{L} alone"
        assert np.array_equal(run_batch(gpu, srv, [(a, b)] + big)[0], want), f"N=2^{L} in a large batch"
    srv.close()


def _trace(torch, n, q, seed, lo, hi, gap_ns):
    rng = np.random.default_rng(seed)
    logs = rng.integers(lo, hi + 1, size=n)
    hosts = [(rand(1 << int(L), q, seed * 1000 + 2 * i), rand(1 << int(L), q, seed * 1000 + 2 * i + 1))
             for i, L in enumerate(logs)]
    dev = [(to_dev(torch, a), to_dev(torch, b)) for a, b in hosts]
    arrival = np.cumsum(rng.exponential(gap_ns, size=n)).astype(np.uint64) if gap_ns else np.zeros(n, np.uint64)
    return logs, hosts, dev, arrival


@pytest.mark.parametrize("name", ["q116", "q122"])
def test_alone_equals_mixed(gpu, orc, name):
    """Each request of a Poisson trace with mixed sizes 2^4 .. 2^14, served
    with continuous batching over several streams, gives the same bits as the
    same request multiplied alone; a sample is checked against the oracle.
    q_116 takes the sparse Montgomery reduction, q122 the dense one."""
    q, _ = POLY_MODULI[name]
    srv = server(q, hbm_budget=1 << 30, max_inflight=4, max_batch=16, min_log2n=4, max_log2n=14)
    logs, hosts, dev, arrival = _trace(gpu, 160, q, 7, 4, 14, 20_000)
    mixed = [gpu.empty_like(a) for a, _ in dev]
    done, st = srv.run_trace([(a, b, c) for (a, b), c in zip(dev, mixed)], arrival)
    assert st["requests"] == 160 and st["batches"] >= 2
    assert np.all(done >= arrival)
    for i, ((a, b), c) in enumerate(zip(dev, mixed)):
        alone = gpu.empty_like(a)
        srv.multiply([(a, b, alone)])
        assert gpu.equal(alone, c), f"request {i} (N=2^{int(logs[i])}) differs alone vs mixed"
    for i in range(0, 160, 16):
        a, b = hosts[i]
        assert np.array_equal(from_dev(mixed[i]), expect(orc, a, b, q)), f"request {i}"
    srv.close()


def test_alone_equals_mixed_large_sizes(gpu, orc):
    """The serving sizes of the benchmark, 2^10 .. 2^16 (two-pass transforms
    from 2^12 on, and 2^10, 2^11 in one or two passes by batch size): each
    request of a Poisson trace equals the same request alone, and a sample
    covering every size matches the oracle's exact NTT pipeline."""
    q = C.Q_116
    srv = server(q, hbm_budget=1 << 30, max_inflight=4, max_batch=16, min_log2n=10, max_log2n=16)
    logs, hosts, dev, arrival = _trace(gpu, 56, q, 21, 10, 16, 30_000)
    mixed = [gpu.empty_like(a) for a, _ in dev]
    done, st = srv.run_trace([(a, b, c) for (a, b), c in zip(dev, mixed)], arrival)
    assert st["requests"] == 56 and st["batches"] >= 2
    for i, ((a, b), c) in enumerate(zip(dev, mixed)):
        alone = gpu.empty_like(a)
        srv.multiply([(a, b, alone)])
        assert gpu.equal(alone, c), f"request {i} (N=2^{int(logs[i])}) differs alone vs mixed"
    for L in range(10, 17):
        i = int(np.flatnonzero(logs == L)[0])
        a, b = hosts[i]
        assert np.array_equal(from_dev(mixed[i]), orc.polymul_ntt(a, b, q)), f"request {i} (N=2^{L})"
    srv.close()


def test_bench_serving_trace_sample(gpu, orc):
    """The benchmark's serving trace as the benchmark builds it (16,384
    requests, N uniform on 2^10 .. 2^16 from seed 55, operands from the
    device generator with seeds 551 and 552, 2 GiB budget, 3 streams, batches
    of 64, every request arrived at t = 0): a seeded sample of 14 requests,
    two of each size, against the oracle's exact NTT pipeline."""
    from benchntt_amd import fixtures as fx
    from benchntt_amd.polymul import PolymulServer

    q = C.Q_116
    n_req = 16384
    logs = np.random.default_rng(55).integers(10, 17, size=n_req)
    sizes = [1 << int(l) for l in logs]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    A, B = fx.uniform(int(offs[-1]), q, 551), fx.uniform(int(offs[-1]), q, 552)
    Cc = gpu.empty_like(A)
    reqs = [(A[offs[i]:offs[i + 1]], B[offs[i]:offs[i + 1]], Cc[offs[i]:offs[i + 1]]) for i in range(n_req)]
    srv = PolymulServer(q, hbm_budget=2 << 30, max_inflight=3, max_batch=64, min_log2n=10, max_log2n=16)
    done, st = srv.run_trace(reqs, np.zeros(n_req, dtype=np.uint64))
    assert st["requests"] == n_req and st["dropped"] == 0
    assert 0 < st["max_batch_ns"] <= int(done.max()) and st["max_loop_gap_ns"] > 0
    gpu.cuda.synchronize()
    rng = np.random.default_rng(1234)
    for L in range(10, 17):
        for i in rng.choice(np.flatnonzero(logs == L), size=2, replace=False):
            a, b, c = (from_dev(t) for t in reqs[int(i)])
            assert np.array_equal(c, orc.polymul_ntt(a, b, q)), f"request {int(i)} (N=2^{L})"
    srv.close()
    del A, B, Cc, reqs


def test_batch_work_cap(gpu):
    """max_batch_butterflies closes a trace batch before a request that would
    take its work above the cap: 40 requests of 2^10..2^13 arriving at once
    under a cap of two 2^13-point requests (3 * 2^12 * 13 butterflies each)
    and a batch limit of 64 take at least as many batches as the work
    requires, and a cap below any one request gives one batch per request;
    every result equals the same request alone."""
    q = C.Q_116
    one13 = 3 * (1 << 12) * 13
    logs, hosts, dev, arrival = _trace(gpu, 40, q, 21, 10, 13, 0)
    work = sum(3 * (1 << (int(L) - 1)) * int(L) for L in logs)
    for cap, min_batches in ((2 * one13, -(-work // (2 * one13))), (1, 40)):
        srv = server(q, hbm_budget=1 << 30, max_inflight=3, max_batch=64, min_log2n=10, max_log2n=13,
                     max_batch_butterflies=cap)
        outs = [gpu.empty_like(a) for a, _ in dev]
        done, st = srv.run_trace([(a, b, c) for (a, b), c in zip(dev, outs)], arrival)
        assert st["batches"] >= min_batches, (cap, st["batches"])
        if cap == 1:
            assert st["batches"] == 40
        for i, ((a, b), c) in enumerate(zip(dev, outs)):
            alone = gpu.empty_like(a)
            srv.multiply([(a, b, alone)])
            assert gpu.equal(alone, c), f"cap {cap} request {i}"
        srv.close()


def test_tight_budget_admission(gpu):
    """A budget of 2.5 requests of 2^12 points: arrived requests wait for
    budget, never more than two are on the device, and results are unchanged."""
    q = C.Q_116
    budget = int(2.5 * req_bytes(12))
    srv = server(q, hbm_budget=budget, max_inflight=4, max_batch=8, min_log2n=12, max_log2n=12)
    logs, hosts, dev, arrival = _trace(gpu, 24, q, 9, 12, 12, 0)
    mixed = [gpu.empty_like(a) for a, _ in dev]
    done, st = srv.run_trace([(a, b, c) for (a, b), c in zip(dev, mixed)], arrival)
    assert st["admission_waits"] > 0
    assert st["dropped"] == 0  # no admission timeout set
    assert 1 <= st["max_inflight_requests"] <= 2
    assert st["batches"] >= 12
    for i, ((a, b), c) in enumerate(zip(dev, mixed)):
        alone = gpu.empty_like(a)
        srv.multiply([(a, b, alone)])
        assert gpu.equal(alone, c), f"request {i}"
    srv.close()


def test_admission_timeout_drops(gpu):
    """Forced from the host: 200 requests of 2^12 points arrive at once under
    a budget of one request in flight and a 2 ms admission timeout.  The
    first request is admitted at once and completes; requests still waiting
    after 2 ms are dropped: their done time is DROPPED and their c keeps the
    value it had; every completed c equals the same request multiplied alone."""
    from benchntt_amd.polymul import DROPPED

    q = C.Q_116
    n_req = 200
    srv = server(q, hbm_budget=int(1.5 * req_bytes(12)), max_inflight=2, max_batch=4, min_log2n=12,
                 max_log2n=12, admit_timeout_ns=2_000_000)
    logs, hosts, dev, arrival = _trace(gpu, n_req, q, 21, 12, 12, 0)
    sentinel = to_dev(gpu, np.full((1 << 12, 2), 7, dtype=np.uint64))
    outs = [sentinel.clone() for _ in range(n_req)]
    done, st = srv.run_trace([(a, b, c) for (a, b), c in zip(dev, outs)], arrival)
    dropped = done == DROPPED
    assert st["requests"] == n_req
    assert st["dropped"] == int(dropped.sum())
    assert 1 <= st["dropped"] < n_req
    assert not dropped[0]
    assert st["max_inflight_requests"] == 1
    for i, ((a, b), c) in enumerate(zip(dev, outs)):
        if dropped[i]:
            assert gpu.equal(c, sentinel), f"dropped request {i} wrote its result"
        else:
            alone = gpu.empty_like(a)
            srv.multiply([(a, b, alone)])
            assert gpu.equal(alone, c), f"request {i}"
    srv.close()


def test_errors(gpu):
    from benchntt_amd import BnttError
    from benchntt_amd.polymul import PolymulServer

    q = C.Q_116
    with pytest.raises(BnttError):
        PolymulServer(q - 1)  # even
    with pytest.raises(BnttError):
        PolymulServer(C.Q_NTT)  # above 2^122
    with pytest.raises(BnttError):
        PolymulServer(97, min_log2n=1, max_log2n=5)  # 2^6 does not divide 96
    with pytest.raises(BnttError):
        PolymulServer(q, max_log2n=17)
    srv = PolymulServer(q, hbm_budget=req_bytes(10), min_log2n=4, max_log2n=12)
    a = to_dev(gpu, rand(1 << 11, q, 1))
    with pytest.raises(BnttError):
        srv.multiply([(a, a, gpu.empty_like(a))])  # larger than the whole budget
    small = to_dev(gpu, rand(8, q, 1))
    with pytest.raises(BnttError):
        srv.multiply([(small, small, gpu.empty_like(small))])  # size not served
    with pytest.raises(ValueError):
        srv.multiply([(a[:100], a[:100], a[:100].clone())])  # not a power of two
    ok = to_dev(gpu, rand(1 << 10, q, 2))
    out = gpu.empty_like(ok)
    other = to_dev(gpu, rand(1 << 10, q, 3))
    with pytest.raises(BnttError):  # one request's result is another's operand
        srv.multiply([(ok, ok, out), (out, other, gpu.empty_like(ok))])
    srv.multiply([(ok, ok, out)])  # the server still works after the errors
    srv.multiply([(other, ok, other)])  # a result may replace its own operand
    srv.close()
