"""GPU parity of the forward NTT against the oracle and the reference's golden
vectors (bit-exact), over the plain one-transform path and uniform batches."""
import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import uniform
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

QN, MUN = C.Q_NTT, C.MU_NTT


def ntt_dev(torch, x, tw, q=QN, mu=MUN, inplace=False):
    import benchntt_amd as bn

    dx, dtw = to_dev(torch, x), to_dev(torch, tw)
    dy = dx if inplace else torch.empty_like(dx)
    bn.ntt(dy, dx, q, mu, dtw)
    torch.cuda.synchronize()
    return from_dev(dy)


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024])
def test_golden_device(gpu, fixtures, n):
    y = ntt_dev(gpu, fixtures[f"x_{n}"], fixtures[f"twd_{n}"])
    assert np.array_equal(y, fixtures[f"ver_{n}"])


@pytest.mark.parametrize("n", [16, 1024])
def test_golden_host_entry(gpu, fixtures, n):
    import benchntt_amd as bn

    x, tw = fixtures[f"x_{n}"].copy(), fixtures[f"twd_{n}"].copy()
    y = np.zeros_like(x)
    bn.ntt(y, x, QN, MUN, tw)
    assert np.array_equal(y, fixtures[f"ver_{n}"])


@pytest.mark.parametrize("L", list(range(1, 19)))
def test_host_entry_sizes(gpu, orc, L):
    """Host-array calls at every size up to 2^18: zero-copy up to 2^12 (the
    kernel posts completion words to pinned memory, which the call waits
    on), staged through pinned memory with the table moved by DMA and x read
    in place up to 2^16 (completion words for the one-launch split at 2^13,
    the stream above), and the runtime's copies above that.  Consecutive calls
    with fresh inputs each match the oracle, the last with a value above q
    (the literal formulas)."""
    import benchntt_amd as bn

    n = 1 << L
    tw = uniform(n, QN, 300 + L)
    calls = 6 if L <= 12 else 3
    for r in range(calls):
        x = uniform(n, QN, 310 + 16 * L + r)
        if r == calls - 1:
            x[0] = [2**64 - 1, 2**64 - 1]
        y = np.zeros_like(x)
        bn.ntt(y, x, QN, MUN, tw)
        assert np.array_equal(y, orc.ntt(x, tw, QN, MUN)), r


def test_host_entry_table_changes(gpu, orc):
    """Consecutive host-array calls whose twiddle tables change between
    calls: the same table twice, another table, the first array modified in
    place (one entry), tables of other sizes and back, and an entry >= 2^124
    (the literal formulas), each call against the oracle."""
    import benchntt_amd as bn

    def call(x, tw):
        y = np.zeros_like(x)
        bn.ntt(y, x, QN, MUN, tw)
        assert np.array_equal(y, orc.ntt(x, tw, QN, MUN))

    n = 1024
    tw1, tw2 = uniform(n, QN, 900), uniform(n, QN, 901)
    x = uniform(n, QN, 902)
    call(x, tw1)
    call(uniform(n, QN, 903), tw1)
    call(x, tw2)
    call(x, tw1)
    tw1[700] = uniform(1, QN, 904)[0]
    call(x, tw1)
    call(uniform(4096, QN, 905), uniform(4096, QN, 906))
    call(uniform(1 << 14, QN, 908), uniform(1 << 14, QN, 909))
    call(uniform(1 << 17, QN, 910), uniform(1 << 17, QN, 911))
    call(uniform(16, QN, 907), tw1[:16].copy())
    call(x, tw1)
    tw1[1000, 1] |= np.uint64(1 << 62)
    call(x, tw1)


@pytest.mark.parametrize("L", list(range(1, 17)))
def test_random_batches(gpu, orc, L):
    """Every size 2^1..2^16, a few transforms per launch (one workgroup may
    hold several transforms, and the tail workgroup is partial)."""
    n = 1 << L
    count = 3 if L <= 12 else 2
    x, tw = uniform(count * n, QN, 10 + L), uniform(n, QN, 40 + L)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L", [10, 11, 12])
@pytest.mark.parametrize("count", [1, 4, 5])
def test_few_transforms(gpu, orc, L, count):
    """One to five transforms of 2^10..2^12 points in one launch, one
    workgroup per transform, against the oracle."""
    n = 1 << L
    x, tw = uniform(count * n, QN, 70 + L + count), uniform(n, QN, 80 + L)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L", [17, 18, 20, 22, 24])
def test_large_plain(gpu, orc, L):
    """Up to the largest size the entry points accept (2^24 points: three
    passes through device memory)."""
    n = 1 << L
    x, tw = uniform(n, QN, L), uniform(n, QN, 2 * L)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L,count", [(13, 1), (16, 1), (19, 1), (21, 1), (14, 160), (13, 512)])
def test_multi_pass_tile_sizes(gpu, orc, L, count):
    """Multi-pass launches pick 512-, 1024- or 2048-value tiles by the values
    in the launch (below 2^21, below 2^22, else), in two or three passes:
    cases of each, against the oracle."""
    n = 1 << L
    x, tw = uniform(count * n, QN, 90 + L), uniform(n, QN, 91 + L)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


def test_uniform_batch_8192x1024(gpu, orc):
    """The uniform-batch workload: 8192 transforms of 1024 points."""
    x, tw = uniform(8192 * 1024, QN, 3), uniform(1024, QN, 4)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L", [4, 10, 12])
def test_in_place(gpu, orc, L):
    n = 1 << L
    x, tw = uniform(5 * n, QN, 1), uniform(n, QN, 2)
    assert np.array_equal(ntt_dev(gpu, x, tw, inplace=True), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L", [10, 13, 17])
def test_host_entry_in_place(gpu, orc, L):
    """A host-array call with Y the same array as X (the reference's ntt()
    reads all of X before it writes Y) on each host path: zero-copy (2^10),
    staged through pinned memory (2^13) and the runtime's copies (2^17)."""
    import benchntt_amd as bn

    n = 1 << L
    x, tw = uniform(n, QN, 970 + L), uniform(n, QN, 971 + L)
    want = orc.ntt(x, tw, QN, MUN)
    bn.ntt(x, x, QN, MUN, tw)
    assert np.array_equal(x, want)


@pytest.mark.parametrize("L", [6, 10, 12])
def test_shifted_overlap_rejected(gpu, L):
    """y = x + N (a shifted overlap) is an argument error for one-pass sizes
    too: one workgroup's stores would land in another transform's input before
    that transform is loaded.  Nothing is launched."""
    import benchntt_amd as bn

    n = 1 << L
    buf = gpu.zeros((6 * n, 2), dtype=gpu.int64, device="cuda")
    tw = gpu.zeros((n, 2), dtype=gpu.int64, device="cuda")
    with pytest.raises(bn.BnttError) as ei:
        bn.ntt(buf[n:5 * n], buf[:4 * n], QN, MUN, tw)
    assert ei.value.code == 1 and "overlap" in str(ei.value)


def test_edge_inputs(gpu, orc):
    """Inputs and twiddles at q, above q and above 2^124 follow the reference
    formulas exactly (the fast reduction hands such operands to the literal
    Barrett path)."""
    n = 256
    rng = np.random.default_rng(0)
    x, tw = uniform(n, QN, 5), uniform(n, QN, 6)
    x[::7, 1] = rng.integers(0, 2**64, size=x[::7].shape[0], dtype=np.uint64)
    tw[::5, 1] |= np.uint64(1 << 60)
    x[3] = [QN & (2**64 - 1), QN >> 64]
    x[4] = [2**64 - 1, 2**64 - 1]
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L", [10, 14])
def test_products_near_zero(gpu, orc, L):
    """The first stage multiplies x[j + N/2] by tw[1]: with x[j + N/2] =
    tw[1]^-1 k (k < 2^20) in three lanes of four, those products are k mod q,
    which takes the fast product's rarely needed final subtraction."""
    import random

    from benchntt_amd.residues import from_ints, to_ints

    n, rng = 1 << L, random.Random(L)
    x, tw = uniform(2 * n, QN, 60 + L), uniform(n, QN, 61 + L)
    inv = pow(to_ints(tw[1:2])[0], -1, QN)
    for t in range(2):
        hi = [inv * rng.randrange(1 << 20) % QN if j % 4 else rng.randrange(QN) for j in range(n // 2)]
        x[t * n + n // 2:(t + 1) * n] = from_ints(hi)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


def test_inputs_just_below_q(gpu, orc):
    """Residues in [q - 2^100, q] (top limb equal to q's) take the fast path,
    whose values then stay at most q; one launch and a batch plan both match
    the oracle, with q itself among the inputs."""
    import random

    from benchntt_amd.batch import BatchPlan
    from benchntt_amd.residues import from_ints

    rng = random.Random(5)
    n = 1024
    x = from_ints([QN - rng.randrange(1 << 100) if j % 3 else QN - (j % 7) for j in range(4 * n)])
    tw = uniform(n, QN, 77)
    want = orc.ntt(x, tw, QN, MUN)
    assert np.array_equal(ntt_dev(gpu, x, tw), want)
    dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
    dy = gpu.empty_like(dx)
    plan = BatchPlan([(dx[t * n:(t + 1) * n], dy[t * n:(t + 1) * n], dtw) for t in range(4)], [], QN, MUN,
                     graph=False)
    plan.execute()
    gpu.cuda.synchronize()
    assert np.array_equal(from_dev(dy), want)
    plan.close()


def test_other_modulus_literal_path(gpu, orc):
    """A modulus outside the fast form runs the literal formulas: one-pass
    and two-pass lone transforms (2^6, 2^10), 2^12 on one workgroup per
    transform and a multi-pass size, two transforms each."""
    q = C.Q_116
    mu = (1 << 235) // q
    for L in (6, 10, 12, 13):
        x, tw = uniform(2 << L, q, L), uniform(1 << L, q, L + 1)
        assert np.array_equal(ntt_dev(gpu, x, tw, q, mu), orc.ntt(x, tw, q, mu))


@pytest.mark.parametrize("L", [8, 10])
def test_lone_pass2_literal(gpu, orc, L):
    """A table entry >= 2^124 that only the second pass reads (entries from
    2^(L/2) on belong to its stages) sends that pass of the wavefronts using it
    to the literal formulas while the first pass stays fast; three transforms
    in one launch against the oracle."""
    n = 1 << L
    x, tw = uniform(3 * n, QN, 800 + L), uniform(n, QN, 801 + L)
    tw[n - 5, 1] |= np.uint64(1 << 61)
    tw[(1 << (L // 2)) + 3, 1] |= np.uint64(1 << 63)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


@pytest.mark.parametrize("L,count", [(8, 64), (8, 65), (10, 64), (10, 65)])
def test_lone_count_boundary(gpu, orc, L, count):
    """Up to 64 transforms of at most 2^10 points run the lone kernel (one
    workgroup each); 65 take the packed one-launch kernel: both sides of the
    boundary against the oracle."""
    n = 1 << L
    x, tw = uniform(count * n, QN, 820 + L + count), uniform(n, QN, 821 + L)
    assert np.array_equal(ntt_dev(gpu, x, tw), orc.ntt(x, tw, QN, MUN))


def test_linearity_2_20(gpu):
    """Size-independent property at 2^20 points: NTT(a + b) = NTT(a) + NTT(b)."""
    import benchntt_amd as bn

    n = 1 << 20
    a, b, tw = uniform(n, QN, 1), uniform(n, QN, 2), uniform(n, QN, 3)
    da, db, dtw = to_dev(gpu, a), to_dev(gpu, b), to_dev(gpu, tw)
    s, fa, fb, fs, t = (gpu.empty_like(da) for _ in range(5))
    bn.vector_add(s, da, db, QN)
    bn.ntt(fa, da, QN, MUN, dtw)
    bn.ntt(fb, db, QN, MUN, dtw)
    bn.ntt(fs, s, QN, MUN, dtw)
    bn.vector_add(t, fa, fb, QN)
    gpu.cuda.synchronize()
    assert gpu.equal(t, fs)


def test_bad_sizes(gpu):
    import benchntt_amd as bn

    x = gpu.zeros((24, 2), dtype=gpu.int64, device="cuda")
    tw = gpu.zeros((16, 2), dtype=gpu.int64, device="cuda")
    with pytest.raises(ValueError):
        bn.ntt(x, x, QN, MUN, tw)  # 24 is not a whole number of 16-point transforms
    tw3 = gpu.zeros((3, 2), dtype=gpu.int64, device="cuda")
    with pytest.raises(ValueError):
        bn.ntt(x[:3], x[:3], QN, MUN, tw3)


def test_c_abi_argument_errors(gpu):
    """The C entry points on bad arguments: count 0 is a successful no-op
    (nothing is read, null pointers included); log2n outside [1, 24], null or
    misaligned pointers and a missing tmp above 4096 points are BNTT_EINVAL
    with a message, and nothing is launched; the host entry rejects sizes
    outside [1, 24] the same way."""
    import ctypes

    from benchntt_amd import _lib

    lib = _lib.load()
    m, u = _lib.U128.of(QN), _lib.U128.of(MUN)
    x = gpu.zeros((1 << 13, 2), dtype=gpu.int64, device="cuda")
    y, tw = gpu.empty_like(x), gpu.zeros_like(x)
    s = ctypes.c_void_p(gpu.cuda.current_stream().cuda_stream)

    def dev(yp, xp, tp, L, count, tmp=None):
        return lib.bntt_ntt_fwd_dev(yp, xp, tp, L, count, ctypes.byref(m), ctypes.byref(u), tmp, s)

    assert dev(None, None, None, 10, 0) == 0
    for L in (0, 25, -3):
        assert dev(y.data_ptr(), x.data_ptr(), tw.data_ptr(), L, 1) == _lib.BNTT_EINVAL
        assert "log2n" in lib.bntt_last_error().decode()
    assert dev(None, x.data_ptr(), tw.data_ptr(), 10, 1) == _lib.BNTT_EINVAL
    assert dev(y.data_ptr(), x.data_ptr() + 8, tw.data_ptr(), 10, 1) == _lib.BNTT_EINVAL
    assert "aligned" in lib.bntt_last_error().decode()
    assert dev(y.data_ptr(), x.data_ptr(), tw.data_ptr(), 13, 1) == _lib.BNTT_EINVAL
    assert "tmp" in lib.bntt_last_error().decode()
    xh = np.zeros((16, 2), np.uint64)
    for L in (0, 25):
        rc = lib.bntt_ntt_fwd(xh.ctypes.data, xh.ctypes.data, xh.ctypes.data, L, ctypes.byref(m), ctypes.byref(u))
        assert rc == _lib.BNTT_EINVAL
    gpu.cuda.synchronize()


@pytest.mark.parametrize("L", [11, 12, 13])
def test_lone_transform_split(gpu, orc, L):
    """A lone transform of 2^11 .. 2^13 points given a tmp (2^L + 16
    residues) splits over 4, 8 or 16 workgroups joined by a grid barrier:
    out of place and three calls in a row on one tmp (each launch's barrier
    token differs from the stale words of the last) against the oracle; up to
    2^12 also in place and the same result without a tmp (one workgroup),
    above that both rejected (the tmp is required and y must not overlap x);
    with an input above q and with a pass-2 twiddle >= 2^124 (literal formulas
    on some wavefronts), over a modulus outside the fast form, under graph
    capture (which takes the unsplit paths: one workgroup, or two passes) with
    replays on refilled inputs, and a tmp overlapping x rejected."""
    import ctypes

    import benchntt_amd as bn
    from benchntt_amd import _lib

    n = 1 << L
    lib = _lib.load()
    need = lib.bntt_ntt_tmp_bytes(L, 1) // 16
    assert need == n + 16
    tmp = gpu.empty((need, 2), dtype=gpu.int64, device="cuda")

    def run(y, xx, dtw, t, q=QN, mu=MUN, stream=None):
        m, u = _lib.U128.of(q), _lib.U128.of(mu)
        _lib.check(lib.bntt_ntt_fwd_dev(y.data_ptr(), xx.data_ptr(), dtw.data_ptr(), L, 1, ctypes.byref(m),
                                        ctypes.byref(u), t.data_ptr() if t is not None else None, stream))

    for seed in range(3):
        x, tw = uniform(n, QN, 500 + 7 * L + seed), uniform(n, QN, 501 + 7 * L + seed)
        want = orc.ntt(x, tw, QN, MUN)
        dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
        dy = gpu.empty_like(dx)
        run(dy, dx, dtw, tmp)
        gpu.cuda.synchronize()
        assert np.array_equal(from_dev(dy), want), seed
    dz = dx.clone()
    dw = gpu.empty_like(dx)
    if L <= 12:
        run(dz, dz, dtw, tmp)
        run(dw, dx, dtw, None)
        gpu.cuda.synchronize()
        assert np.array_equal(from_dev(dz), want)
        assert np.array_equal(from_dev(dw), want)
    else:
        with pytest.raises(bn.BnttError):
            run(dz, dz, dtw, tmp)
        with pytest.raises(bn.BnttError):
            run(dw, dx, dtw, None)

    # literal formulas: an input above q (pass 1 of one wavefront, then pass 2
    # of every wavefront that reads its outputs) and a table entry >= 2^124
    # that only the last stage reads
    xe, twe = x.copy(), tw.copy()
    xe[5] = [2**64 - 1, 2**64 - 1]
    twe[n - 3, 1] |= np.uint64(1 << 62)
    dye = gpu.empty_like(dx)
    run(dye, to_dev(gpu, xe), to_dev(gpu, twe), tmp)
    gpu.cuda.synchronize()
    assert np.array_equal(from_dev(dye), orc.ntt(xe, twe, QN, MUN))
    q, mu = C.Q_116, (1 << 235) // C.Q_116
    xq, twq = uniform(n, q, 600 + L), uniform(n, q, 601 + L)
    dyq = gpu.empty_like(dx)
    run(dyq, to_dev(gpu, xq), to_dev(gpu, twq), tmp, q, mu)
    gpu.cuda.synchronize()
    assert np.array_equal(from_dev(dyq), orc.ntt(xq, twq, q, mu))

    # captured: replays follow refilled inputs
    s = gpu.cuda.Stream()
    dg, dyg = gpu.empty_like(dx), gpu.empty_like(dx)
    dg.copy_(dx)
    gpu.cuda.synchronize()
    g = gpu.cuda.CUDAGraph()
    with gpu.cuda.graph(g, stream=s):
        run(dyg, dg, dtw, tmp, stream=ctypes.c_void_p(s.cuda_stream))
    for r in range(3):
        xr = uniform(n, QN, 700 + 5 * L + r)
        dg.copy_(to_dev(gpu, xr))
        gpu.cuda.synchronize()
        g.replay()
        gpu.cuda.synchronize()
        assert np.array_equal(from_dev(dyg), orc.ntt(xr, tw, QN, MUN)), r
    del g

    buf = gpu.empty((3 * n, 2), dtype=gpu.int64, device="cuda")
    buf[:n] = dx
    with pytest.raises(bn.BnttError) as ei:
        run(buf[n:2 * n], buf[:n], dtw, buf[n // 2:n // 2 + need])
    assert "tmp" in str(ei.value)
    # the tmp's barrier words after its intermediate count as part of it
    big = gpu.empty((2 * n + 16, 2), dtype=gpu.int64, device="cuda")
    big[n + 8:2 * n + 8] = dx
    with pytest.raises(bn.BnttError):
        run(gpu.empty_like(dx), big[n + 8:2 * n + 8], dtw, big[:need])


@pytest.mark.parametrize("L", [12, 13])
def test_python_as0_sizes(gpu, orc, L):
    """bn.ntt's as0, the reference's scratch argument: one transform with an
    as0 of the reference's N residues (too small for the split's N + 16: the
    wrapper allocates its own) or of N + 16 (used), and several transforms of
    2^13 points with an as0 shorter than count * N rejected."""
    import benchntt_amd as bn

    n = 1 << L
    x, tw = uniform(n, QN, 960 + L), uniform(n, QN, 961 + L)
    want = orc.ntt(x, tw, QN, MUN)
    dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
    for extra in (0, 16):
        as0 = gpu.zeros((n + extra, 2), dtype=gpu.int64, device="cuda")
        dy = gpu.empty_like(dx)
        bn.ntt(dy, dx, QN, MUN, dtw, as0=as0)
        gpu.cuda.synchronize()
        assert np.array_equal(from_dev(dy), want), extra
        assert bool((as0 != 0).any()) == (extra == 16), extra
    if L > 12:
        x2 = gpu.cat([dx, dx])
        with pytest.raises(ValueError):
            bn.ntt(gpu.empty_like(x2), x2, QN, MUN, dtw, as0=gpu.empty((n, 2), dtype=gpu.int64, device="cuda"))


@pytest.mark.parametrize("L", [12, 13])
def test_lone_split_concurrent_streams(gpu, orc, L):
    """Lone 2^12 or 2^13 transforms split over 8 or 16 workgroups, enqueued
    on three streams at once, each stream with its own tmp (own barrier words
    and tokens): every stream's results match the oracle, over several
    rounds."""
    import ctypes

    from benchntt_amd import _lib

    n = 1 << L
    lib = _lib.load()
    need = lib.bntt_ntt_tmp_bytes(L, 1) // 16
    m, mu = _lib.U128.of(QN), _lib.U128.of(MUN)
    streams = [gpu.cuda.Stream() for _ in range(3)]
    tmps = [gpu.empty((need, 2), dtype=gpu.int64, device="cuda") for _ in streams]
    tw = uniform(n, QN, 950)
    dtw = to_dev(gpu, tw)
    xs = [uniform(n, QN, 951 + i) for i in range(3)]
    dxs = [to_dev(gpu, x) for x in xs]
    wants = [orc.ntt(x, tw, QN, MUN) for x in xs]
    gpu.cuda.synchronize()
    for rnd in range(4):
        ys = [gpu.empty_like(dx) for dx in dxs]
        for _ in range(5):
            for s, t, dx, dy in zip(streams, tmps, dxs, ys):
                _lib.check(lib.bntt_ntt_fwd_dev(dy.data_ptr(), dx.data_ptr(), dtw.data_ptr(), L, 1, ctypes.byref(m),
                                                ctypes.byref(mu), t.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        gpu.cuda.synchronize()
        for i, (dy, want) in enumerate(zip(ys, wants)):
            assert np.array_equal(from_dev(dy), want), (rnd, i)

