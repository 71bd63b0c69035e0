"""GPU parity of uneven batches (persistent work-queue kernel, graph replay)
against the oracle, transform by transform and product by product."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import uniform
from benchntt_amd.workloads import uneven_pairs
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

QN, MUN = C.Q_NTT, C.MU_NTT


def build(torch, wl, own_tables=False):
    """Device buffers, transform tasks and the pair products of `wl`.
    own_tables: every transform reads its own device copy of its size's
    table, so no two transforms share a table pointer and no product can be
    fused into its transforms' last pass."""
    dx = to_dev(torch, wl.x)
    dy = torch.empty_like(dx)
    dc = torch.empty((wl.pair_total, 2), dtype=torch.int64, device="cuda")
    tabs = {l: to_dev(torch, t) for l, t in wl.tables.items()}
    tasks, muls = [], []
    for i, (l, off) in enumerate(zip(wl.log2n, wl.offsets)):
        n = 1 << int(l)
        tw = tabs[int(l)].clone() if own_tables else tabs[int(l)]
        tasks.append((dx[off:off + n], dy[off:off + n], tw))
    for j, poff in enumerate(wl.pair_offsets):
        n = 1 << int(wl.log2n[2 * j])
        a0, b0 = wl.offsets[2 * j], wl.offsets[2 * j + 1]
        muls.append((dy[a0:a0 + n], dy[b0:b0 + n], dc[poff:poff + n]))
    return dx, dy, dc, tabs, tasks, muls


def check_all(orc, wl, y, c, threads=16, every=1, q=QN, mu=MUN):
    """Every transform (or every `every`-th pair) and its product vs the oracle."""
    def one(j):
        outs = []
        for i in (2 * j, 2 * j + 1):
            l, off = int(wl.log2n[i]), int(wl.offsets[i])
            n = 1 << l
            want = orc.ntt(wl.x[off:off + n], wl.tables[l], q, mu)
            if not np.array_equal(y[off:off + n], want):
                return f"transform {i} (N=2^{l}) differs"
            outs.append(want)
        poff = int(wl.pair_offsets[j])
        n = outs[0].shape[0]
        if not np.array_equal(c[poff:poff + n], orc.vvmul(outs[0], outs[1], q, mu)):
            return f"product {j} differs"
        return None

    with ThreadPoolExecutor(threads) as ex:
        errs = [e for e in ex.map(one, range(0, len(wl.pair_offsets), every)) if e]
    assert not errs, errs[:5]


@pytest.mark.parametrize("tw_mode", ["global", "lds", "wide"])
def test_small_uneven_batch(gpu, orc, tw_mode):
    """Sizes 2^1..2^16 with their products, for each placement of the
    twiddles (staged in LDS within the three- or two-workgroup budget, or read
    from device memory)."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(80, QN, seed=11, lo=1, hi=16)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=False, tw_mode=tw_mode)
    plan.execute()
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc))
    info = plan.info()
    # every product is fused into its pair's last pass: the two passes of the
    # two-pass sizes are the only levels
    assert info["levels"] == 2 and info["items"] > 0
    plan.close()
    # the same without products
    plan = BatchPlan(tasks, [], QN, MUN, graph=False, tw_mode=tw_mode)
    assert plan.info()["levels"] == 2
    plan.close()


@pytest.mark.parametrize("tw_mode", ["global", "lds"])
def test_unfused_products(gpu, orc, tw_mode):
    """Transforms that each read their own table cannot share tiles: their
    products run as product items after their operands' last pass (two-pass
    transforms in two start groups: passes at levels 0-1 and 1-2, products
    after each group's last pass at levels 2 and 3)."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(80, QN, seed=12, lo=1, hi=16)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl, own_tables=True)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=False, tw_mode=tw_mode)
    plan.execute()
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc))
    assert plan.info()["levels"] == 4
    plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_fixed_tables(gpu, orc, graph):
    """Plans whose tables are converted once, at creation, over several
    executions with the inputs refilled between them; a table with an entry
    >= 2^124 is detected at creation and its transforms (and their fused
    products) take the literal launch."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(60, QN, seed=19, lo=3, hi=15)
    sizes = sorted({int(l) for l in wl.log2n})
    wl.tables[sizes[-2]][5, 1] |= np.uint64(1 << 62)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=graph, fixed_tables=True)
    stream = gpu.cuda.Stream()
    for rep in range(3):
        if rep:
            wl.x = uniform(wl.total, QN, 300 + rep)
            dx.copy_(to_dev(gpu, wl.x))
            gpu.cuda.synchronize()
        with gpu.cuda.stream(stream):
            plan.execute()
        stream.synchronize()
        check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()


def test_graph_replays_follow_inputs(gpu, orc):
    """Replays of the captured graph, with eager work and a refill of the
    inputs in between, each match the oracle (counters rearm every launch)."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(40, QN, seed=5, lo=4, hi=14)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=True)
    stream = gpu.cuda.Stream()
    for rep in range(3):
        if rep:
            wl.x = uniform(wl.total, QN, 100 + rep)
            dx.copy_(to_dev(gpu, wl.x))
            gpu.cuda.synchronize()
        with gpu.cuda.stream(stream):
            plan.execute()
        stream.synchronize()
        check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()


@pytest.mark.parametrize("tune", [
    {l: (1, 1) for l in range(1, 10)} | {l: (2, 2) for l in range(10, 15)},
    {l: (2, 1) for l in range(2, 11)} | {l: (3, 3) for l in range(11, 17)},
    {l: (3, 2) for l in range(6, 17)},
])
@pytest.mark.parametrize("tw_mode", ["global", "wide"])
def test_tuning_choices_agree(gpu, orc, tune, tw_mode):
    """Every radix / pass-count choice gives the reference's outputs."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(60, QN, seed=7, lo=1, hi=16)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, tune=tune, graph=False, tw_mode=tw_mode)
    plan.execute()
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()


def test_large_sizes_three_passes(gpu, orc):
    from benchntt_amd.batch import BatchPlan

    tasks, xs, tws, ys = [], [], [], []
    for l in (17, 18, 20, 24):
        x, tw = uniform(1 << l, QN, l), uniform(1 << l, QN, 50 + l)
        dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
        dy = gpu.empty_like(dx)
        tasks.append((dx, dy, dtw))
        xs.append(x), tws.append(tw), ys.append(dy)
    plan = BatchPlan(tasks, [], QN, MUN, graph=False, tw_mode="wide")
    plan.execute()
    gpu.cuda.synchronize()
    for x, tw, dy in zip(xs, tws, ys):
        assert np.array_equal(from_dev(dy), orc.ntt(x, tw, QN, MUN))
    plan.close()


def test_edge_values_in_batch(gpu, orc):
    """Transforms holding values >= q (and >= 2^124) take the literal path
    inside the batch kernel without disturbing their neighbours."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(24, QN, seed=3, lo=4, hi=13)
    wl.x[5::97, 1] = np.uint64(0xFFFFFFFFFFFFFFFF)
    wl.x[7::131] = [QN & (2**64 - 1), QN >> 64]
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    for mode in ("global", "lds", "wide"):
        plan = BatchPlan(tasks, muls, QN, MUN, graph=False, tw_mode=mode)
        plan.execute()
        gpu.cuda.synchronize()
        check_all(orc, wl, from_dev(dy), from_dev(dc))
        plan.close()


def test_literal_modulus(gpu, orc):
    """A modulus outside the fast form 2^124 - c (q_116 with the reference's
    MBITS=116 constant) runs every level on the literal formulas."""
    from benchntt_amd.batch import BatchPlan

    q = C.Q_116
    mu = (1 << 235) // q
    wl = uneven_pairs(40, q, seed=13, lo=1, hi=14)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, q, mu, graph=False)
    plan.execute()
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc), q=q, mu=mu)
    plan.close()


def test_flagged_table_and_replays(gpu, orc):
    """A twiddle table with an entry >= 2^124 sends every transform using it
    (and the products of their outputs) to the literal launch after each
    level; graph replays that alternate inputs with and without values >=
    2^124 each match the oracle (the punt lists are cleared every execution)."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(40, QN, seed=17, lo=4, hi=15)
    sizes = sorted({int(l) for l in wl.log2n})
    wl.tables[sizes[-1]][3, 1] = np.uint64(0xFFFFFFFFFFFFFFFF)   # a two-pass size
    wl.tables[sizes[1]][1, 1] |= np.uint64(1 << 61)              # a one-pass size
    clean = wl.x.copy()
    dirty = wl.x.copy()
    dirty[11::193, 1] = np.uint64(0xFFFFFFFFFFFFFFFF)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=True)
    stream = gpu.cuda.Stream()
    for rep in range(4):
        wl.x = dirty if rep % 2 == 0 else clean
        dx.copy_(to_dev(gpu, wl.x))
        gpu.cuda.synchronize()
        with gpu.cuda.stream(stream):
            plan.execute()
        stream.synchronize()
        check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()


def test_launch_each_matches(gpu, orc):
    from benchntt_amd.batch import launch_each

    wl = uneven_pairs(30, QN, seed=9, lo=4, hi=15)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    tmp = gpu.empty((1 << 16, 2), dtype=gpu.int64, device="cuda")
    launch_each(tasks, QN, MUN, tmp)
    gpu.cuda.synchronize()
    y = from_dev(dy)
    for i, (l, off) in enumerate(zip(wl.log2n, wl.offsets)):
        n = 1 << int(l)
        assert np.array_equal(y[off:off + n], orc.ntt(wl.x[off:off + n], wl.tables[int(l)], QN, MUN)), i


def test_uneven_20000_full_check(gpu, orc):
    """The uneven-batch workload at full size: 20,000 transforms with N
    log-uniform on 2^4..2^16 in pairs (about 2^27.6 coefficients), each pair's
    product after its transforms; every transform and product bit-exact."""
    from benchntt_amd.batch import BatchPlan, load_tune_table

    wl = uneven_pairs(20000, QN, seed=2024)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    # as the benchmark runs it: tuned, graph replay, tables converted at creation
    plan = BatchPlan(tasks, muls, QN, MUN, tune=load_tune_table(), graph=True, fixed_tables=True)
    stream = gpu.cuda.Stream()
    with gpu.cuda.stream(stream):
        plan.execute()
    stream.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc), threads=16)
    plan.close()


@pytest.mark.parametrize("L", [10, 4, 5, 6, 7, 8, 9])
def test_uniform_bench_shapes(gpu, orc, L):
    """The uniform batches the benchmark times, at full size and with its
    seeds: 8192 x N=1024 and 2^23 coefficients of each N = 16..512, one shared
    table, as a tuned plan with fixed tables replayed from a graph (two
    executions); every output against the oracle."""
    from benchntt_amd.batch import BatchPlan, load_tune_table

    total, n = 1 << 23, 1 << L
    cnt = total // n
    x, tw = uniform(total, QN, 5), uniform(n, QN, 40 + L)
    dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
    dy = gpu.empty_like(dx)
    tasks = [(dx[i * n:(i + 1) * n], dy[i * n:(i + 1) * n], dtw) for i in range(cnt)]
    stream = gpu.cuda.Stream()
    with gpu.cuda.stream(stream):
        plan = BatchPlan(tasks, [], QN, MUN, tune=load_tune_table(), graph=True, fixed_tables=True)
        plan.execute()
        plan.execute()
    stream.synchronize()
    assert np.array_equal(from_dev(dy), orc.ntt(x, tw, QN, MUN))
    plan.close()


def test_product_on_part_of_an_output(gpu, orc):
    """A product reading the middle of a two-pass transform's output (not its
    start) runs after that transform's last pass."""
    from benchntt_amd.batch import BatchPlan

    L = 16
    x, tw = uniform(1 << L, QN, 90), uniform(1 << L, QN, 91)
    x2, tw2 = uniform(1 << L, QN, 92), uniform(1 << L, QN, 93)
    dx, dtw, dx2, dtw2 = (to_dev(gpu, a) for a in (x, tw, x2, tw2))
    dy, dy2 = gpu.empty_like(dx), gpu.empty_like(dx2)
    dc = gpu.empty((4096, 2), dtype=gpu.int64, device="cuda")
    plan = BatchPlan([(dx, dy, dtw), (dx2, dy2, dtw2)], [(dy[5000:9096], dy2[12345:16441], dc)], QN, MUN,
                     graph=False)
    plan.execute()
    gpu.cuda.synchronize()
    y, y2 = orc.ntt(x, tw, QN, MUN), orc.ntt(x2, tw2, QN, MUN)
    assert np.array_equal(from_dev(dc), orc.vvmul(y[5000:9096], y2[12345:16441], QN, MUN))
    plan.close()


def test_plan_rejects_overlapping_tasks(gpu):
    from benchntt_amd import BnttError
    from benchntt_amd.batch import BatchPlan

    a = to_dev(gpu, uniform(4096, QN, 1))
    tw = to_dev(gpu, uniform(1024, QN, 2))
    out = gpu.empty((4096, 2), dtype=gpu.int64, device="cuda")
    with pytest.raises(BnttError):  # the second transform writes what the first reads
        BatchPlan([(a[:1024], out[:1024], tw), (a[1024:2048], a[512:1536], tw)], [], QN, MUN)
    with pytest.raises(BnttError):  # two transforms write overlapping outputs
        BatchPlan([(a[:1024], out[:1024], tw), (a[1024:2048], out[512:1536], tw)], [], QN, MUN)
    with pytest.raises(BnttError):  # a product writes a transform's input
        BatchPlan([(a[:1024], out[:1024], tw)], [(out[:1024], out[:1024], a[:1024])], QN, MUN)
    with pytest.raises(BnttError):  # a product writes another product's operand
        BatchPlan([], [(a[:512], a[:512], out[:512]), (out[256:768], a[:512], out[1024:1536])], QN, MUN)
    # in place, and products reading outputs: accepted
    BatchPlan([(a[:1024], a[:1024], tw), (a[1024:2048], out[:1024], tw)],
              [(out[:1024], a[:1024], out[2048:3072])], QN, MUN).close()


@pytest.mark.parametrize("plan_graph", [False, True])
def test_capture_in_caller_graph(gpu, orc, plan_graph):
    """plan.execute() captured in the caller's own graph: every replay, with
    the inputs refilled eagerly between replays, matches the oracle."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(24, QN, seed=21, lo=4, hi=14)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=plan_graph)
    s = gpu.cuda.Stream()
    s.wait_stream(gpu.cuda.current_stream())
    g = gpu.cuda.CUDAGraph()
    with gpu.cuda.graph(g, stream=s):
        plan.execute()
    for rep in range(3):
        wl.x[:] = uniform(wl.x.shape[0], QN, 300 + rep)
        dx.copy_(to_dev(gpu, wl.x))
        gpu.cuda.synchronize()
        g.replay()
        gpu.cuda.synchronize()
        check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()


def test_contiguous_packed_items(gpu, orc):
    """Single-pass transforms back to back in memory are read through one
    base pointer per item; a gap or a permutation falls back to per-transform
    pointers.  Both against the oracle."""
    from benchntt_amd.batch import BatchPlan

    for L, cnt in ((6, 300), (10, 37), (11, 9)):
        n = 1 << L
        x = uniform(cnt * n + n, QN, 500 + L)
        tw = uniform(n, QN, 600 + L)
        dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
        dy = gpu.empty_like(dx)
        # back to back, then one with a gap of one transform, then reversed order
        layouts = [list(range(cnt)), list(range(cnt // 2)) + list(range(cnt // 2 + 1, cnt + 1)),
                   list(reversed(range(cnt)))]
        for lay in layouts:
            tasks = [(dx[i * n:(i + 1) * n], dy[i * n:(i + 1) * n], dtw) for i in lay]
            plan = BatchPlan(tasks, [], QN, MUN, graph=False)
            plan.execute()
            gpu.cuda.synchronize()
            y = from_dev(dy)
            for i in lay[:: max(1, cnt // 7)] + [lay[-1]]:
                assert np.array_equal(y[i * n:(i + 1) * n], orc.ntt(x[i * n:(i + 1) * n], tw, QN, MUN)), (L, i)
            plan.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_plans(gpu, orc, seed):
    """Seeded random plans: 1..300 transforms of sizes 2^1..2^14 whose tables
    are shared (a table of 2^14 also serves every smaller size: a transform of
    2^L reads only entries below 2^L, and the plan converts one table once
    for all its users), some transforms in place, products on
    random pairs of equal size (a transform may feed several), executed from
    a graph or not, each twiddle placement; every output against the oracle."""
    from benchntt_amd.batch import BatchPlan

    rng = np.random.default_rng(1000 + seed)
    n_t = int(rng.integers(1, 301))
    logs = [int(v) for v in rng.integers(1, 15, size=n_t)]
    big = uniform(1 << 14, QN, 7000 + seed)
    own = {L: uniform(1 << L, QN, 7100 + 20 * seed + L) for L in set(logs)}
    dbig = to_dev(gpu, big)
    down = {L: to_dev(gpu, t) for L, t in own.items()}
    shared = [bool(rng.integers(0, 2)) for _ in logs]
    inplace = [L <= 11 and bool(rng.integers(0, 3) == 0) for L in logs]
    xs = [uniform(1 << L, QN, 8000 + 500 * seed + i) for i, L in enumerate(logs)]
    dxs = [to_dev(gpu, x) for x in xs]
    dys = [dx if ip else gpu.empty_like(dx) for dx, ip in zip(dxs, inplace)]
    # a shared table is passed as its prefix view: the same device pointer
    tasks = [(dx, dy, dbig[:1 << L] if sh else down[L]) for dx, dy, sh, L in zip(dxs, dys, shared, logs)]
    by_size = {}
    for i, L in enumerate(logs):
        by_size.setdefault(L, []).append(i)
    pairs = []
    for L, idx in by_size.items():
        for _ in range(len(idx) // 2 + 1):
            i, j = (int(v) for v in rng.choice(idx, size=2))
            pairs.append((i, j))
    dcs = [gpu.empty_like(dys[i]) for i, _ in pairs]
    muls = [(dys[i], dys[j], dc) for (i, j), dc in zip(pairs, dcs)]
    graph = seed % 2 == 0
    tw_mode = ("global", "lds", "wide")[seed % 3]
    print(f"\nrandom plan seed {seed}: {n_t} transforms, {len(pairs)} products, graph {graph}, {tw_mode}", flush=True)
    plan = BatchPlan(tasks, muls, QN, MUN, graph=graph, tw_mode=tw_mode)
    plan.execute()
    gpu.cuda.synchronize()
    want = [orc.ntt(x, (big if sh else own[L])[:1 << L], QN, MUN) for x, sh, L in zip(xs, shared, logs)]
    for i, dy in enumerate(dys):
        assert np.array_equal(from_dev(dy), want[i]), f"transform {i} (N=2^{logs[i]})"
    for k, ((i, j), dc) in enumerate(zip(pairs, dcs)):
        assert np.array_equal(from_dev(dc), orc.vvmul(want[i], want[j], QN, MUN)), f"product {k}"
    plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_executions_on_two_streams_in_order(gpu, orc, graph):
    """Executions of one plan enqueued on two streams without host
    synchronisation run one after the other (the second stream waits for the
    first execution): the shared work-queue counters stay consistent, so a
    later execution after the outputs are cleared is still complete."""
    from benchntt_amd.batch import BatchPlan

    wl = uneven_pairs(60, QN, seed=77, lo=4, hi=14)
    dx, dy, dc, tabs, tasks, muls = build(gpu, wl)
    s1, s2 = gpu.cuda.Stream(), gpu.cuda.Stream()
    with gpu.cuda.stream(s1):
        plan = BatchPlan(tasks, muls, QN, MUN, graph=graph)
    gpu.cuda.synchronize()
    for _ in range(3):
        plan.execute(s1)
        plan.execute(s2)
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc))
    dy.zero_()
    dc.zero_()
    gpu.cuda.synchronize()
    plan.execute(s2)
    gpu.cuda.synchronize()
    check_all(orc, wl, from_dev(dy), from_dev(dc))
    plan.close()
