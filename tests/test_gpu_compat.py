"""The drop-in on the GPU: a C caller written like the reference's drivers
(text fixtures read with fscanf "%lu" into hi/lo words, converted to
uint128_t, then ntt() / vector_add() / vector_sub() / vector_mul() / axpy()
with the reference's arguments) is compiled against include/benchntt_compat.h,
linked with libbntt.so and run; its outputs must equal the expected files."""
import os
import subprocess

import numpy as np
import pytest

from benchntt_amd import config as C

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r"""
#define stages %(stages)d
#define N (1 << stages)
#include <inttypes.h>
#include "benchntt_compat.h"

/* hi, lo decimal words per residue, as the reference's fixtures */
static int load(const char *path, uint128_t *v, int n)
{
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    for (int i = 0; i < n; i++) {
        uint64_t hi, lo;
        if (fscanf(f, "%%" SCNu64 " %%" SCNu64, &hi, &lo) != 2) { fclose(f); return 1; }
        v[i] = INT128(hi, lo);
    }
    fclose(f);
    return 0;
}

static int same(const uint128_t *a, const uint128_t *b, int n)
{
    for (int i = 0; i < n; i++)
        if (a[i] != b[i]) return 0;
    return 1;
}

int main(int argc, char **argv)
{
    const char *d = argv[1];
    char p[512];
    static uint128_t x[N], y[N], tw[N], ver[N], as0[N], as1[N];
    uint128_t modulus = INT128(%(qn_hi)dULL, %(qn_lo)dULL), mu = INT128(%(mun_hi)dULL, %(mun_lo)dULL);
    snprintf(p, sizeof p, "%%s/x_%%d.txt", d, N); if (load(p, x, N)) return 2;
    snprintf(p, sizeof p, "%%s/twd_%%d.txt", d, N); if (load(p, tw, N)) return 2;
    snprintf(p, sizeof p, "%%s/ver_%%d.txt", d, N); if (load(p, ver, N)) return 2;
    ntt(y, x, modulus, mu, tw, as0, as1);
    printf("ntt %%d %%s\n", N, same(y, ver, N) ? "ok" : "MISMATCH");
#if stages == 10
    static uint128_t a[1024], b[1024], c[1024], want[1024];
    const char *ops[4] = {"vvadd", "vvsub", "vvmul", "axpy"};
    uint128_t s = INT128(0x02E4B338A15CE95FULL, 0xB1C09DAA097FEBD4ULL);
    snprintf(p, sizeof p, "%%s/twd_1024.txt", d); if (load(p, a, 1024)) return 2;
    snprintf(p, sizeof p, "%%s/x_1024.txt", d); if (load(p, b, 1024)) return 2;
    for (int k = 0; k < 4; k++) {
        snprintf(p, sizeof p, "%%s/%%s_1024.txt", d, ops[k]); if (load(p, want, 1024)) return 2;
        if (k == 0) vector_add(c, a, b, MODULUS);
        if (k == 1) vector_sub(c, a, b, MODULUS);
        if (k == 2) vector_mul(c, a, b, MODULUS, MU);
        if (k == 3) axpy(c, a, b, s, MODULUS, MU);
        printf("%%s %%s\n", ops[k], same(c, want, 1024) ? "ok" : "MISMATCH");
    }
    /* host primitives against the device result of one product */
    printf("mulmod128 %%s\n", mulmod128(a[5], b[5], MODULUS, MU) == (vector_mul(c, a, b, MODULUS, MU), c[5]) ? "ok" : "MISMATCH");
#endif
    return 0;
}
"""


@pytest.mark.parametrize("stages", [4, 7, 10])
def test_reference_style_caller(gpu, fixtures, tmp_path, stages):
    from benchntt_amd import fixtures as fx

    data = tmp_path / "data"
    data.mkdir()
    for k, v in fixtures.items():
        fx.write_text(str(data / f"{k}.txt"), v, final_newline=not k.startswith(("x_", "twd_", "ver_")))
    m = (1 << 64) - 1
    src = tmp_path / "caller.c"
    src.write_text(CALLER % {"stages": stages, "qn_hi": C.Q_NTT >> 64, "qn_lo": C.Q_NTT & m,
                             "mun_hi": C.MU_NTT >> 64, "mun_lo": C.MU_NTT & m})
    exe = tmp_path / "caller"
    lib_dir = os.path.join(ROOT, "benchntt_amd")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", lib_dir, "-lbntt", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    print(f"\nrunning the reference-style caller, N = 2^{stages}", flush=True)
    r = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    assert f"ntt {1 << stages} ok" in lines, r.stdout
    if stages == 10:
        for op in ("vvadd", "vvsub", "vvmul", "axpy", "mulmod128"):
            assert f"{op} ok" in lines, r.stdout


@pytest.mark.parametrize("n", [16, 1024])
def test_calltime_ntt_output(gpu, fixtures, n):
    """The reference-protocol timing harness (calls timed in C) leaves the
    golden output of its last host-array ntt() call in Y."""
    from benchntt_amd import calltime

    x, tw = fixtures[f"x_{n}"].copy(), fixtures[f"twd_{n}"].copy()
    y = np.zeros_like(x)
    ns = calltime.ntt_ns_per_call(y, x, tw, C.Q_NTT, C.MU_NTT, iters=6)
    assert ns > 0
    assert np.array_equal(y, fixtures[f"ver_{n}"])


@pytest.mark.parametrize("op", ["vvadd", "vvsub", "vvmul", "axpy"])
def test_calltime_blas_output(gpu, fixtures, op):
    from benchntt_amd import calltime

    a, b = fixtures["twd_1024"].copy(), fixtures["x_1024"].copy()
    c = np.zeros_like(a)
    ns = calltime.blas_ns_per_call(op, c, a, b, C.Q_BLAS, C.MU_BLAS, C.AXPY_SCALAR, iters=6)
    assert ns > 0
    assert np.array_equal(c, fixtures[f"{op}_1024"])
