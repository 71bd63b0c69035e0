"""GPU parity of the polynomial-multiply server's transforms on their own
(bntt_pm_transform): the forward negacyclic NTT against the oracle's exact
forward network with the negacyclic table, the inverse against the oracle's
exact inverse network, closed forms fixed by the mathematics (the transform of
1 and of x), round trips and the product identity at 2^16 points, in place,
launches split by the HBM budget, and argument errors."""
import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints, to_ints
from conftest import from_dev, to_dev
from oracle import negacyclic_tables, two_adic_root
from test_gpu_serve import rand, server
from test_oracle import POLY_MODULI

pytestmark = pytest.mark.gpu


def _brv(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def _psi(q, L):
    w, v = two_adic_root(q)
    return pow(w, 1 << (v - L - 1), q)


def _blocks(count, n, q, seed):
    return np.concatenate([rand(n, q, seed + t) for t in range(count)])


@pytest.mark.parametrize("name", sorted(POLY_MODULI))
def test_forward_inverse_oracle(gpu, orc, name):
    """Sizes 2^1 .. 2^12 (up to the modulus' 2-adicity), three transforms per
    call: forward = the exact network with the negacyclic table, inverse = the
    exact inverse network with the inverse table, transform by transform."""
    q, lmax = POLY_MODULI[name]
    top = min(lmax, 12)
    srv = server(q, min_log2n=1, max_log2n=top, max_batch=8)
    for L in range(1, top + 1):
        n, count = 1 << L, 3
        tw, itw = negacyclic_tables(q, L)
        x = _blocks(count, n, q, 50 * L)
        dx = to_dev(gpu, x)
        dy = gpu.empty_like(dx)
        srv.forward(dy, dx, n)
        y = from_dev(dy)
        dz = gpu.empty_like(dx)
        srv.inverse(dz, dy, n)
        z = from_dev(dz)
        for t in range(count):
            blk = slice(t * n, (t + 1) * n)
            fy = orc.ntt_exact(x[blk], from_ints(tw), q)
            assert np.array_equal(y[blk], fy), f"forward N=2^{L} transform {t}"
            assert np.array_equal(z[blk], orc.intt_exact(fy, from_ints(itw), q)), f"inverse N=2^{L} transform {t}"
            assert np.array_equal(z[blk], x[blk]), f"round trip N=2^{L} transform {t}"
    srv.close()


@pytest.mark.parametrize("L", [13, 14, 16])
def test_two_pass_sizes_oracle(gpu, orc, L):
    """Two-pass sizes over q_116 against the oracle's exact networks."""
    q, n = C.Q_116, 1 << L
    tw, itw = negacyclic_tables(q, L)
    srv = server(q, min_log2n=L, max_log2n=L, max_batch=4)
    x = _blocks(2, n, q, 700 + L)
    dx = to_dev(gpu, x)
    dy = gpu.empty_like(dx)
    srv.forward(dy, dx, n)
    y = from_dev(dy)
    fy = orc.ntt_exact(x[:n], from_ints(tw), q)
    assert np.array_equal(y[:n], fy)
    dz = gpu.empty_like(dx)
    srv.inverse(dz, dy, n)
    assert np.array_equal(from_dev(dz)[:n], orc.intt_exact(fy, from_ints(itw), q))
    assert np.array_equal(from_dev(dz), x)
    srv.close()


@pytest.mark.parametrize("name", ["q116", "q122", "q64"])
def test_closed_forms(gpu, name):
    """F(1) = (1, .., 1); F(x)[i] = psi^(2 brv(i) + 1) for the root psi of
    order 2N; the inverse of (1, .., 1) is 1."""
    q, _ = POLY_MODULI[name]
    srv = server(q, min_log2n=3, max_log2n=16, max_batch=8)
    for L in (3, 9, 11, 12, 16):
        n = 1 << L
        psi = _psi(q, L)
        assert pow(psi, n, q) == q - 1
        one, xpoly = [0] * n, [0] * n
        one[0], xpoly[1] = 1, 1
        dx = to_dev(gpu, np.concatenate([from_ints(one), from_ints(xpoly)]))
        dy = gpu.empty_like(dx)
        srv.forward(dy, dx, n)
        y = to_ints(from_dev(dy))
        assert y[:n] == [1] * n, f"F(1) N=2^{L}"
        assert y[n:] == [pow(psi, 2 * _brv(i, L) + 1, q) for i in range(n)], f"F(x) N=2^{L}"
        dz = gpu.empty_like(dx)
        srv.inverse(dz, dy, n)
        assert to_ints(from_dev(dz)) == one + xpoly, f"inverse N=2^{L}"
    srv.close()


def test_round_trip_in_place_2_16(gpu):
    """64 transforms of 2^16 points (64 MiB), forward then inverse in place,
    give the input back; the result is the same computed out of place."""
    q, L = C.Q_116, 16
    n = 1 << L
    srv = server(q, min_log2n=L, max_log2n=L, max_batch=32, hbm_budget=1 << 30)
    from benchntt_amd import fixtures as fx

    x = fx.uniform(64 * n, q, 11)
    ref = x.clone()
    out = gpu.empty_like(x)
    srv.forward(out, x, n)
    srv.forward(x, x, n)
    assert gpu.equal(x, out)
    srv.inverse(x, x, n)
    assert gpu.equal(x, ref)
    srv.close()


def test_product_identity(gpu):
    """inverse(F(a) * F(b) mod q) = the server's product a * b, at 2^12 and
    2^16 points (the pointwise products taken on the host)."""
    q = C.Q_116
    srv = server(q, min_log2n=12, max_log2n=16, max_batch=4)
    for L in (12, 16):
        n = 1 << L
        a, b = rand(n, q, 900 + L), rand(n, q, 950 + L)
        da, db = to_dev(gpu, a), to_dev(gpu, b)
        fa, fb = gpu.empty_like(da), gpu.empty_like(db)
        srv.forward(fa, da, n)
        srv.forward(fb, db, n)
        prod = from_ints([u * v % q for u, v in zip(to_ints(from_dev(fa)), to_ints(from_dev(fb)))])
        dp = to_dev(gpu, prod)
        srv.inverse(dp, dp, n)
        c = gpu.empty_like(da)
        srv.multiply([(da, db, c)])
        assert gpu.equal(dp, c), f"N=2^{L}"
    srv.close()


def test_budget_and_batch_split(gpu, orc):
    """Ten two-pass transforms with room for three pass buffers, and ten
    one-pass transforms with max_batch 1 (three per launch): each call runs
    several launches and matches the transforms computed one by one."""
    q = C.Q_116
    for L, kw in ((12, {"hbm_budget": 3 * 16 << 12, "max_batch": 8}), (10, {"max_batch": 1})):
        n = 1 << L
        srv = server(q, min_log2n=L, max_log2n=L, **kw)
        x = _blocks(10, n, q, 40 + L)
        dx = to_dev(gpu, x)
        dy = gpu.empty_like(dx)
        srv.forward(dy, dx, n)
        one = gpu.empty_like(dx[:n])
        for t in (0, 4, 9):
            srv.forward(one, dx[t * n:(t + 1) * n], n)
            assert gpu.equal(one, dy[t * n:(t + 1) * n]), f"N=2^{L} transform {t}"
        dz = gpu.empty_like(dx)
        srv.inverse(dz, dy, n)
        assert gpu.equal(dz, dx)
        srv.close()


def test_errors(gpu):
    from benchntt_amd import BnttError, _lib

    q = C.Q_116
    srv = server(q, min_log2n=4, max_log2n=12, hbm_budget=(16 << 12) - 256)
    x = to_dev(gpu, rand(1 << 12, q, 1))
    with pytest.raises(BnttError):
        srv.forward(gpu.empty_like(x), x, 1 << 12)  # a pass buffer exceeds the budget
    with pytest.raises(BnttError):
        srv.forward(gpu.empty_like(x[:8]), x[:8], 8)  # size not served
    with pytest.raises(ValueError):
        srv.forward(gpu.empty_like(x[:96]), x[:96], 64)  # not whole transforms
    with pytest.raises(BnttError):
        srv.forward(x[16:16 + 1024], x[:1024], 1024)  # partial overlap
    with pytest.raises(BnttError):
        _lib.call("bntt_pm_transform", srv._h, 2, x.data_ptr(), x.data_ptr(), 10, 1)  # direction
    y = gpu.empty_like(x[:1024])
    srv.forward(y, x[:1024], 1024)  # the server still works after the errors
    srv.inverse(y, y, 1024)
    assert gpu.equal(y, x[:1024])
    srv.close()
