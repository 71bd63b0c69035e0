"""GPU parity of the input side: seeded uniform residues and root-of-unity
twiddle tables generated on the device against the oracle, word-order
conversion kernels, and their argument errors."""
import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints
from conftest import from_dev, to_dev
from test_io import GEN_MODULI
from test_oracle import POLY_MODULI

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q", GEN_MODULI)
def test_uniform_matches_oracle(gpu, orc, q):
    from benchntt_amd import fixtures as fx

    for n, seed in ((1, 0), (1000, 5), (1 << 18, 77), (100003, 2**64 - 1)):
        print(f"\nuniform q={q:#x} n={n} seed={seed}", flush=True)
        want = orc.uniform(n, q, seed)
        assert np.array_equal(from_dev(fx.uniform(n, q, seed)), want)
        if n <= 1000:
            assert np.array_equal(fx.uniform(n, q, seed, device=False), want)


def test_uniform_large(gpu, orc):
    """2^24 residues (a BLAS vector): prefix and a strided sample against the
    oracle, and all below q."""
    from benchntt_amd import fixtures as fx

    n, q = 1 << 24, C.Q_BLAS
    t = fx.uniform(n, q, 42)
    a = from_dev(t)
    want = orc.uniform(n, q, 42)
    assert np.array_equal(a, want)


@pytest.mark.parametrize("name", ["q116", "q122", "q64", "q97"])
def test_twiddles_match_oracle(gpu, name):
    import oracle
    from benchntt_amd import fixtures as fx

    q, lmax = POLY_MODULI[name]
    for kind in ("cyclic", "negacyclic", "cyclic_inv", "negacyclic_inv"):
        top = lmax + 1 if kind.startswith("cyclic") else lmax
        for L in sorted({1, 2, 3, min(top, 11), top}):
            print(f"\ntwiddles {name} {kind} L={L}", flush=True)
            want = from_ints(oracle.root_table(q, L, kind))
            assert np.array_equal(from_dev(fx.twiddles(L, q, kind)), want), (kind, L)
            if L <= 8:
                assert np.array_equal(fx.twiddles(L, q, kind, device=False), want)


def test_twiddles_errors(gpu):
    from benchntt_amd import BnttError
    from benchntt_amd import fixtures as fx

    with pytest.raises(BnttError):
        fx.twiddles(4, C.Q_NTT, "cyclic")  # >= 2^122
    with pytest.raises(BnttError):
        fx.twiddles(4, 91, "cyclic")  # 7 * 13
    with pytest.raises(BnttError):
        fx.twiddles(17, C.Q_116, "negacyclic")  # needs 2^18 | q - 1
    with pytest.raises(BnttError):
        fx.twiddles(25, C.Q_116, "cyclic")
    fx.twiddles(17, C.Q_116, "cyclic")  # 2^17 | q - 1: fine


def test_word_order_kernels(gpu):
    from benchntt_amd import fixtures as fx

    rng = np.random.default_rng(8)
    for n in (1, 7, 1000, (1 << 20) + 3):
        a = rng.integers(0, 1 << 64, size=(n, 2), dtype=np.uint64)
        d = to_dev(gpu, a)
        hl = fx.u128_to_hilo(d)
        assert np.array_equal(from_dev(hl), a[:, ::-1])
        assert gpu.equal(fx.hilo_to_u128(hl), d)
        fx.hilo_to_u128(d, out=d)  # in place
        assert np.array_equal(from_dev(d), a[:, ::-1])
    from benchntt_amd import BnttError

    x = to_dev(gpu, rng.integers(0, 1 << 64, size=(9, 2), dtype=np.uint64))
    with pytest.raises(BnttError):
        fx.hilo_to_u128(x[:8], out=x[1:])  # shifted overlap


def test_golden_through_device_conversion(gpu, fixtures):
    """The golden NTT case with its input and twiddles uploaded in the
    fixture files' hi/lo order and converted on the device."""
    import benchntt_amd as bn
    from benchntt_amd import fixtures as fx

    x = to_dev(gpu, fixtures["x_1024"][:, ::-1].copy())
    tw = to_dev(gpu, fixtures["twd_1024"][:, ::-1].copy())
    fx.hilo_to_u128(x, out=x)
    fx.hilo_to_u128(tw, out=tw)
    y = gpu.empty_like(x)
    bn.ntt(y, x, C.Q_NTT, C.MU_NTT, tw)
    assert np.array_equal(from_dev(y), fixtures["ver_1024"])
