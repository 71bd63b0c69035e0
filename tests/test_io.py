"""CPU checks of the input side: the oracle's seeded generator and root
tables (pinned by published splitmix64 outputs, exact orders of the roots and
the DFT closed form), the library's host primitives and word-order
converters, and the reference's text fixture format."""
import os
import random
import subprocess

import numpy as np
import pytest

from benchntt_amd import config as C
from benchntt_amd.residues import from_ints, to_ints
from test_oracle import POLY_MODULI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DATA = "/root/reference/data"

Q64 = (1 << 64) - (1 << 32) + 1
GEN_MODULI = [C.Q_NTT, C.Q_BLAS, C.Q_116, Q64, 97, (1 << 128) - 159, 3]


def test_splitmix_known_answers(orc):
    """First outputs of splitmix64 from state 0, as published with the
    generator (Vigna's splitmix64.c test output; Java SplittableRandom)."""
    assert [orc.splitmix_at(0, j) for j in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                         0x06C45D188009454F]


@pytest.mark.parametrize("q", GEN_MODULI)
def test_uniform_in_range_and_balanced(orc, q):
    n = 1 << 15
    v = to_ints(orc.uniform(n, q, 1234))
    assert all(0 <= x < q for x in v)
    if q > 1 << 20:
        top = sum(1 for x in v if x >= q // 2) / n  # about half above q/2
        assert abs(top - 0.5) < 0.02
        assert len(set(v)) == n
    # a different seed gives different values; the same seed the same
    assert to_ints(orc.uniform(64, q, 1234)) != to_ints(orc.uniform(64, q, 1235)) or q < 5
    assert np.array_equal(orc.uniform(100, q, 7), orc.uniform(100, q, 7))


def test_uniform_first_draw_is_splitmix_words(orc):
    """For q_N almost every first draw is accepted: residue i is then the
    masked words 2i (low) and 2i+1 (high) of the sequence."""
    n = 1000
    v = to_ints(orc.uniform(n, C.Q_NTT, 99))
    for i in range(n):
        w = ((orc.splitmix_at(99, 2 * i + 1) << 64) | orc.splitmix_at(99, 2 * i)) & ((1 << 124) - 1)
        assert w >= C.Q_NTT or v[i] == w


@pytest.mark.parametrize("name", ["q116", "q122", "q64", "q97"])
def test_root_tables(orc, name):
    """Stage roots have exactly the stated order, *_inv tables are entry-wise
    inverses, and with the tables the exact network is the cyclic DFT /
    negacyclic NTT in bit-reversed order (O(N^2) evaluation)."""
    import oracle

    q, lmax = POLY_MODULI[name]
    for L in sorted({1, 2, min(lmax, 5)}):
        n = 1 << L
        tabs = {k: oracle.root_table(q, L, k) for k in ("cyclic", "negacyclic", "cyclic_inv", "negacyclic_inv")}
        for s in range(L):
            wc, wn = tabs["cyclic"][(1 << s) + (1 << (s - 1) if s else 0)], tabs["negacyclic"][1 << s]
            # cyclic: entry k = 2^(s-1) (brv_s(k) = 1) holds w_s itself, of order
            # 2^(s+1); stage 0's only entry is w_0^0 = 1; negacyclic entry k = 0 holds p_s
            if s:
                assert pow(wc, 1 << s, q) == q - 1
            else:
                assert wc == 1
            assert pow(wn, 1 << (s + 1), q) == q - 1
        for fk, ik in (("cyclic", "cyclic_inv"), ("negacyclic", "negacyclic_inv")):
            assert all(a * b % q == 1 for a, b in zip(tabs[fk][1:], tabs[ik][1:]))
        x = [random.Random(L).randrange(q) for _ in range(n)]
        brv = [int(format(i, f"0{L}b")[::-1], 2) for i in range(n)]
        yc = to_ints(orc.ntt_exact(from_ints(x), from_ints(tabs["cyclic"]), q)) if q >= 1 << 64 else None
        w = tabs["cyclic"][(1 << (L - 1)) + (1 << (L - 2))] if L > 1 else q - 1  # order 2^L
        psi = tabs["negacyclic"][1 << (L - 1)]  # order 2^(L+1)
        if yc is not None:
            ref = [sum(x[j] * pow(w, j * k, q) for j in range(n)) % q for k in range(n)]
            assert yc == [ref[brv[i]] for i in range(n)]
            yn = to_ints(orc.ntt_exact(from_ints(x), from_ints(tabs["negacyclic"]), q))
            ref = [sum(x[j] * pow(psi, j * (2 * k + 1), q) for j in range(n)) % q for k in range(n)]
            assert yn == [ref[brv[i]] for i in range(n)]


def _edge_operands(q):
    vals = [0, 1, 2, q - 1, q - 2, q, q + 1, (1 << 124) - 1, (1 << 127) + 5, (1 << 128) - 1, 1 << 64,
            (1 << 64) - 1, q >> 1]
    rng = random.Random(q)
    vals += [rng.randrange(q) for _ in range(40)]
    return [v for v in vals if v < 1 << 128]


def test_host_primitives_match_oracle(orc):
    """bntt_addmod128 / submod128 / mulmod128 (library, host) against the
    oracle's restatement on random and edge operands, including the addmod
    edge and operands >= q and >= 2^124."""
    from benchntt_amd import fixtures as fx

    for q, mu in ((C.Q_NTT, C.MU_NTT), (C.Q_BLAS, C.MU_BLAS)):
        ops = _edge_operands(q)
        ml = q & ((1 << 64) - 1)
        ops += [((q >> 64) << 64 | ml) - 5, 5]  # a + b with lo64 = lo64(q)
        for a in ops:
            for b in ops[::3]:
                assert fx.addmod128(a, b, q) == orc.addmod(a, b, q), (a, b)
                assert fx.submod128(a, b, q) == orc.submod(a, b, q), (a, b)
                assert fx.mulmod128(a, b, q, mu) == orc.mulmod(a, b, q, mu), (a, b)


def test_hilo_host_converters():
    from benchntt_amd import fixtures as fx

    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 1000):
        a = rng.integers(0, 1 << 64, size=(n, 2), dtype=np.uint64)
        hl = fx.u128_to_hilo(a)
        assert np.array_equal(hl, a[:, ::-1])
        assert np.array_equal(fx.hilo_to_u128(hl), a)
    out = np.empty((999, 2), dtype=np.uint64)
    with pytest.raises(ValueError):
        fx.u128_to_hilo(a, out=out)  # shorter than the input: refused before any write
    with pytest.raises(ValueError):
        fx.twiddles(4, C.Q_116, "cyclic-inverse", device=False)


def test_text_format_roundtrip(tmp_path, fixtures):
    """write_text / read_text round trip on the golden arrays; with the
    reference's data directory present, the written text is byte-identical to
    its files (the BLAS outputs end with a newline, the NTT files do not)."""
    from benchntt_amd import fixtures as fx

    for name in ("x_16", "twd_1024", "ver_256", "axpy_1024"):
        p = tmp_path / f"{name}.txt"
        fx.write_text(str(p), fixtures[name], final_newline=name.startswith("axpy"))
        assert np.array_equal(fx.read_text(str(p)), fixtures[name])
        ref = os.path.join(REF_DATA, f"{name}.txt")
        if os.path.exists(ref):
            assert p.read_bytes() == open(ref, "rb").read()


def test_fixture_sets_and_tiling(tmp_path, fixtures):
    """load_ntt_set / load_blas_set read a reference-layout data directory;
    N > 1024 repeats the 1024-point input and table (config.c:35-44, 101-134)."""
    from benchntt_amd import fixtures as fx

    for k, v in fixtures.items():
        fx.write_text(str(tmp_path / f"{k}.txt"), v)
    s = fx.load_ntt_set(str(tmp_path), 64)
    assert np.array_equal(s["x"], fixtures["x_64"]) and np.array_equal(s["ver"], fixtures["ver_64"])
    big = fx.load_ntt_set(str(tmp_path), 4096)
    assert big["ver"] is None
    assert big["x"].shape == (4096, 2)
    for j in range(4):
        assert np.array_equal(big["x"][1024 * j:1024 * (j + 1)], fixtures["x_1024"])
        assert np.array_equal(big["tw"][1024 * j:1024 * (j + 1)], fixtures["twd_1024"])
    b = fx.load_blas_set(str(tmp_path))
    assert np.array_equal(b["a"], fixtures["twd_1024"]) and np.array_equal(b["vvmul"], fixtures["vvmul_1024"])


def test_compat_primitives_run(tmp_path, orc):
    """A C caller of the compat header's addmod128 / submod128 / mulmod128
    (host functions: no GPU needed) prints the oracle's values."""
    src = tmp_path / "prim.c"
    src.write_text(
        '#include "benchntt_compat.h"\n'
        "#include <inttypes.h>\n"
        "int main(void) {\n"
        "  uint128_t m = MODULUS, mu = MU;\n"
        "  uint128_t a = m - 3, b = INT128(0x0123456789ABCDEFULL, 0xFEDCBA9876543210ULL);\n"
        "  uint128_t r[3] = {addmod128(a, b, m), submod128(b, a, m), mulmod128(a, b, m, mu)};\n"
        '  for (int i = 0; i < 3; i++) printf("%016" PRIx64 "%016" PRIx64 "\\n", HI64(r[i]), LO64(r[i]));\n'
        "  return 0;\n}\n")
    lib_dir = os.path.join(ROOT, "benchntt_amd")
    exe = tmp_path / "prim"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", lib_dir, "-lbntt", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    m, mu = C.Q_BLAS, C.MU_BLAS
    a, b = m - 3, (0x0123456789ABCDEF << 64) | 0xFEDCBA9876543210
    want = [orc.addmod(a, b, m), orc.submod(b, a, m), orc.mulmod(a, b, m, mu)]
    assert [int(x, 16) for x in out] == want
