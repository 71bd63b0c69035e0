"""CPU checks of the C ABI library: it loads, exports exactly what
include/bntt.h declares, and fails loudly (no CPU fallback) without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if not name.endswith(".h") or name == "benchntt_compat.h":
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        syms |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bntt_[a-z0-9_]+)\s*\(", text, re.M))
    return syms


def test_exports_match_header():
    from benchntt_amd import _lib

    lib_path = _lib.LIB_PATH
    assert os.path.exists(lib_path)
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    decl = declared_symbols()
    assert decl, "no declarations found"
    assert decl <= exported, decl - exported
    assert {s for s in exported if s.startswith("bntt_")} == decl
    assert set(_lib.EXPORTS) == decl


def test_loads_and_versions():
    from benchntt_amd import _lib

    lib = _lib.load()
    assert lib.bntt_version() == 100
    assert lib.bntt_ntt_tmp_bytes(12, 5) == 0
    assert lib.bntt_ntt_tmp_bytes(10, 1) == 0
    # optional: with it a lone 2^11 or 2^12 transform splits over workgroups
    # (intermediate plus 16 residues of grid-barrier words); required from
    # 2^13 on, where one transform of 2^13 points splits the same way
    for L in (11, 12, 13):
        assert lib.bntt_ntt_tmp_bytes(L, 1) == ((1 << L) + 16) * 16
    assert lib.bntt_ntt_tmp_bytes(14, 1) == (1 << 14) * 16
    assert lib.bntt_ntt_tmp_bytes(13, 2) == 2 * (1 << 13) * 16
    assert lib.bntt_ntt_tmp_bytes(11, 2) == 0
    assert lib.bntt_ntt_tmp_bytes(12, 2) == 0
    assert lib.bntt_ntt_tmp_bytes(16, 3) == 3 * (1 << 16) * 16
    assert lib.bntt_ntt_tmp_bytes(24, 1 << 40) == 2**64 - 1  # does not fit a size_t


def test_no_cpu_fallback_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import benchntt_amd as bn

    a = np.zeros((16, 2), np.uint64)
    with pytest.raises(bn.BnttError):
        bn.vector_mul(a, a, a, bn.config.Q_BLAS, bn.config.MU_BLAS)
    with pytest.raises(bn.BnttError):
        bn.ntt(a, a, bn.config.Q_NTT, bn.config.MU_NTT, a)


def test_calltime_library_loads():
    """The timing harness (the reference's per-call protocol in C) links to
    libbntt.so and exports its two entry points; without a GPU a timed call
    reports the library's error instead of a time."""
    import torch

    from benchntt_amd import calltime

    out = subprocess.run(["nm", "-D", "--defined-only", calltime.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported == {"bntt_calltime_ntt", "bntt_calltime_blas"}
    calltime.load()
    with pytest.raises(ValueError):
        calltime.ntt_ns_per_call(np.zeros((16, 2), np.uint64), np.zeros((8, 2), np.uint64),
                                 np.zeros((16, 2), np.uint64), 17, 1)
    if torch.cuda.is_available():
        return
    import benchntt_amd as bn

    a = np.zeros((16, 2), np.uint64)
    with pytest.raises(bn.BnttError):
        calltime.ntt_ns_per_call(a.copy(), a, a, bn.config.Q_NTT, bn.config.MU_NTT, iters=4)
    with pytest.raises(bn.BnttError):
        calltime.blas_ns_per_call("axpy", a.copy(), a, a, bn.config.Q_BLAS, bn.config.MU_BLAS, 3, iters=4)


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes structure of the Python binding has the size and field
    offsets of its C struct in include/bntt.h (compiled with gcc)."""
    from benchntt_amd import _lib

    structs = {"bntt_u128": _lib.U128, "bntt_ntt_task": _lib.NttTask, "bntt_mul_task": _lib.MulTask,
               "bntt_tune_entry": _lib.TuneEntry, "bntt_batch_info": _lib.BatchInfo,
               "bntt_pm_request": _lib.PmRequest, "bntt_pm_config": _lib.PmConfig, "bntt_pm_stats": _lib.PmStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "bntt.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = []
    for cname, py in structs.items():
        want.append(f"{cname} size {ctypes.sizeof(py)}")
        want += [f"{cname} {fname} {getattr(py, fname).offset}" for fname, _ in py._fields_]
    assert got[:-1] == want


def test_compat_header_compiles(tmp_path):
    """include/benchntt_compat.h gives C callers the reference's own call
    shapes (ntt, vector_add, ...); a translation unit written like the
    reference's drivers compiles and links against libbntt.so."""
    src = tmp_path / "caller.c"
    src.write_text(
        "#define stages 10\n"
        "#define N (1 << stages)\n"
        '#include "benchntt_compat.h"\n'
        "int main(void) {\n"
        "  static uint128_t x[N], y[N], tw[N], as0[N], as1[N], c[1024];\n"
        "  uint128_t m = MODULUS, mu = MU;\n"
        "  ntt(y, x, m, mu, tw, as0, as1);\n"
        "  vector_add(c, x, tw, m); vector_sub(c, x, tw, m);\n"
        "  vector_mul(c, x, tw, m, mu); axpy(c, x, tw, m, m, mu);\n"
        "  return 0;\n}\n"
    )
    lib_dir = os.path.join(ROOT, "benchntt_amd")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", lib_dir, "-lbntt", "-Wl,-rpath," + lib_dir, "-o", str(tmp_path / "caller")],
                   check=True)


def test_read_only_outputs_rejected():
    """Host-array calls check their arguments before any library call: a
    read-only output array is an argument error (no GPU needed)."""
    import numpy as np

    import benchntt_amd as bn
    from benchntt_amd import config as C

    x = np.zeros((16, 2), dtype=np.uint64)
    y = np.zeros_like(x)
    y.flags.writeable = False
    with pytest.raises(ValueError, match="writeable"):
        bn.ntt(y, x, C.Q_NTT, C.MU_NTT, x.copy())
    with pytest.raises(ValueError, match="writeable"):
        bn.vector_add(y, x, x, C.Q_BLAS)
