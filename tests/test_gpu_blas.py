"""GPU parity of vvadd / vvsub / vvmul / axpy against the oracle and the
reference's golden vectors (bit-exact)."""
import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import uniform
from conftest import from_dev, to_dev
from test_u128_host import operands

pytestmark = pytest.mark.gpu

QB, MUB, S = C.Q_BLAS, C.MU_BLAS, C.AXPY_SCALAR


def run_dev(torch, op, a, b, x=S, q=QB, mu=MUB):
    import benchntt_amd as bn

    da, db = to_dev(torch, a), to_dev(torch, b)
    dc = torch.empty_like(da)
    if op == "vvadd":
        bn.vector_add(dc, da, db, q)
    elif op == "vvsub":
        bn.vector_sub(dc, da, db, q)
    elif op == "vvmul":
        bn.vector_mul(dc, da, db, q, mu)
    else:
        bn.axpy(dc, da, db, x, q, mu)
    torch.cuda.synchronize()
    return from_dev(dc)


def want(orc, op, a, b, x=S, q=QB, mu=MUB):
    if op == "vvadd":
        return orc.vvadd(a, b, q)
    if op == "vvsub":
        return orc.vvsub(a, b, q)
    if op == "vvmul":
        return orc.vvmul(a, b, q, mu)
    return orc.axpy(a, b, x, q, mu)


OPS = ["vvadd", "vvsub", "vvmul", "axpy"]


@pytest.mark.parametrize("op", OPS)
def test_golden_device(gpu, fixtures, op):
    a, b = fixtures["twd_1024"], fixtures["x_1024"]
    assert np.array_equal(run_dev(gpu, op, a, b), fixtures[f"{op}_1024"])


@pytest.mark.parametrize("op", OPS)
def test_golden_host_entry(gpu, fixtures, op):
    import benchntt_amd as bn

    a, b = fixtures["twd_1024"].copy(), fixtures["x_1024"].copy()
    c = np.zeros_like(a)
    fn = {"vvadd": lambda: bn.vector_add(c, a, b, QB), "vvsub": lambda: bn.vector_sub(c, a, b, QB),
          "vvmul": lambda: bn.vector_mul(c, a, b, QB, MUB), "axpy": lambda: bn.axpy(c, a, b, S, QB, MUB)}[op]
    fn()
    assert np.array_equal(c, fixtures[f"{op}_1024"])


@pytest.mark.parametrize("n", [1, 3, 255, 256, 511, 513, 1000, 65537])
@pytest.mark.parametrize("op", OPS)
def test_ragged_sizes(gpu, orc, op, n):
    a, b = uniform(n, QB, n), uniform(n, QB, n + 7)
    assert np.array_equal(run_dev(gpu, op, a, b), want(orc, op, a, b))


@pytest.mark.parametrize("op", OPS)
def test_edge_operands(gpu, orc, op):
    """Operands at and above q and 2^124, and the reference addmod's edge."""
    a, b = operands(QB, 23)
    assert np.array_equal(run_dev(gpu, op, a, b), want(orc, op, a, b))


@pytest.mark.parametrize("q,mu", [(C.Q_BLAS, C.MU_BLAS), (C.Q_NTT, C.MU_NTT)], ids=["q_blas", "q_ntt"])
def test_products_near_zero(gpu, orc, q, mu):
    """Three lanes in four hold pairs whose product is k < 2^20 mod q (b =
    a^-1 k; for axpy b = s^-1 k): the fast product's folded value is then k
    or k + q, so its rarely needed final subtraction runs, in waves mixed
    with ordinary lanes."""
    import random

    from benchntt_amd.residues import from_ints

    rng = random.Random(3)
    n = 16384
    av, ks = [rng.randrange(1, q) for _ in range(n)], []
    for i in range(n):
        ks.append(rng.randrange(1 << 20) if i % 4 else rng.randrange(q))
    a = from_ints(av)
    b = from_ints([pow(x, -1, q) * k % q for x, k in zip(av, ks)])
    assert np.array_equal(run_dev(gpu, "vvmul", a, b, q=q, mu=mu), want(orc, "vvmul", a, b, q=q, mu=mu))
    s = S % q
    b2 = from_ints([pow(s, -1, q) * k % q for k in ks])
    assert np.array_equal(run_dev(gpu, "axpy", a, b2, x=s, q=q, mu=mu), want(orc, "axpy", a, b2, x=s, q=q, mu=mu))


@pytest.mark.parametrize("op", ["vvadd", "vvsub", "vvmul", "axpy"])
def test_other_modulus_literal_path(gpu, orc, op):
    """A modulus outside the fast form (q_116 with the reference's unused
    MBITS=116 constant) takes the literal reference formulas."""
    q = C.Q_116
    mu = (1 << 235) // q
    a, b = uniform(4099, q, 1), uniform(4099, q, 2)
    x = 12345678901234567890123
    assert np.array_equal(run_dev(gpu, op, a, b, x, q, mu), want(orc, op, a, b, x, q, mu))


def test_full_size_2_24(gpu, orc):
    """The benchmark size: 2^24 residues per vector, every op bit-exact."""
    n = 1 << 24
    a, b = uniform(n, QB, 2024), uniform(n, QB, 2025)
    for op in OPS:
        assert np.array_equal(run_dev(gpu, op, a, b), want(orc, op, a, b)), op


def test_zero_length(gpu):
    import benchntt_amd as bn

    e = gpu.empty((0, 2), dtype=gpu.int64, device="cuda")
    bn.vector_add(e, e, e, QB)
    bn.axpy(e, e, e, S, QB, MUB)
    gpu.cuda.synchronize()


def test_errors(gpu):
    import benchntt_amd as bn

    a = gpu.zeros((8, 2), dtype=gpu.int64, device="cuda")
    b = gpu.zeros((9, 2), dtype=gpu.int64, device="cuda")
    with pytest.raises(ValueError):
        bn.vector_add(a, a, b, QB)
    raw = gpu.zeros(8 * 16 + 8, dtype=gpu.uint8, device="cuda")
    mis = raw[8:].view(gpu.int64)  # 8-byte aligned, not 16
    with pytest.raises(bn.BnttError):
        bn.vector_add(mis, mis, mis, QB)


@pytest.mark.parametrize("op", OPS)
def test_in_place_and_shifted_overlap(gpu, orc, op):
    """c may be a or b (the result replaces that operand); c shifted by one
    residue against an operand is rejected before any launch."""
    import benchntt_amd as bn

    n = 1000
    a, b = uniform(n, QB, 91), uniform(n, QB, 92)
    for target in ("a", "b"):
        da, db = to_dev(gpu, a), to_dev(gpu, b)
        dc = da if target == "a" else db
        call = {"vvadd": lambda: bn.vector_add(dc, da, db, QB), "vvsub": lambda: bn.vector_sub(dc, da, db, QB),
                "vvmul": lambda: bn.vector_mul(dc, da, db, QB, MUB),
                "axpy": lambda: bn.axpy(dc, da, db, S, QB, MUB)}[op]
        call()
        gpu.cuda.synchronize()
        assert np.array_equal(from_dev(dc), want(orc, op, a, b)), target
    x = to_dev(gpu, uniform(n + 1, QB, 93))
    with pytest.raises(bn.BnttError):
        bn.vector_add(x[1:], x[:n], x[:n], QB)
    with pytest.raises(bn.BnttError):
        bn.axpy(x[:n], to_dev(gpu, a), x[1:], S, QB, MUB)


def test_host_entry_points_from_many_threads(gpu, orc):
    """The host-pointer entry points are thread-safe: 8 threads call axpy,
    vvmul and ntt on host arrays at once (shared staging buffers), each result
    matches the oracle."""
    from concurrent.futures import ThreadPoolExecutor

    import benchntt_amd as bn

    def work(i):
        n = 1000 + 517 * i
        a, b = uniform(n, C.Q_BLAS, 40 + i), uniform(n, C.Q_BLAS, 80 + i)
        c = np.empty_like(a)
        bn.axpy(c, a, b, C.AXPY_SCALAR, C.Q_BLAS, C.MU_BLAS)
        if not np.array_equal(c, orc.axpy(a, b, C.AXPY_SCALAR, C.Q_BLAS, C.MU_BLAS)):
            return f"axpy thread {i}"
        bn.vector_mul(c, a, b, C.Q_BLAS, C.MU_BLAS)
        if not np.array_equal(c, orc.vvmul(a, b, C.Q_BLAS, C.MU_BLAS)):
            return f"vvmul thread {i}"
        L = 8 + (i % 7)
        x, tw = uniform(1 << L, C.Q_NTT, 120 + i), uniform(1 << L, C.Q_NTT, 160 + i)
        y = np.empty_like(x)
        bn.ntt(y, x, C.Q_NTT, C.MU_NTT, tw)
        if not np.array_equal(y, orc.ntt(x, tw, C.Q_NTT, C.MU_NTT)):
            return f"ntt thread {i}"
        return None

    with ThreadPoolExecutor(8) as ex:
        errs = [e for r in range(3) for e in ex.map(work, range(8 * r, 8 * r + 8)) if e]
    assert not errs, errs
