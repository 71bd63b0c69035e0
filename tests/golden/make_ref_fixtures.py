"""Convert the reference's golden data files into tests/golden/ref_fixtures.npz.

The reference holds its only known-answer vectors as text under data/: one
decimal u64 per line, the high word then the low word of each 128-bit element
(format read by config.c:6-197):
  x_N.txt, twd_N.txt, ver_N.txt   N = 16..1024  (ntt_scalar.c:661-701)
  vvadd/vvsub/vvmul/axpy_1024.txt (blas_scalar.c:618-749; inputs a = twd_1024,
                                   b = x_1024)
This script reads those data files as text (it runs no reference code) and
stores every array as uint64 (n, 2) in the native layout used everywhere in
this repository: column 0 = low word, column 1 = high word.

Run from the repository root:  python tests/golden/make_ref_fixtures.py
"""
import os
import sys

import numpy as np

REF_DATA = os.environ.get("BNTT_REFERENCE_DATA", "/root/reference/data")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_fixtures.npz")


def read_hilo(path: str) -> np.ndarray:
    words = np.loadtxt(path, dtype=np.uint64).reshape(-1, 2)  # (hi, lo) rows
    return np.ascontiguousarray(words[:, ::-1])  # -> (lo, hi)


def main() -> int:
    if not os.path.isdir(REF_DATA):
        print(f"no reference data at {REF_DATA}", file=sys.stderr)
        return 1
    arrays = {}
    for log2n in range(4, 11):
        n = 1 << log2n
        for kind in ("x", "twd", "ver"):
            a = read_hilo(os.path.join(REF_DATA, f"{kind}_{n}.txt"))
            assert a.shape == (n, 2), (kind, n, a.shape)
            arrays[f"{kind}_{n}"] = a
    for op in ("vvadd", "vvsub", "vvmul", "axpy"):
        a = read_hilo(os.path.join(REF_DATA, f"{op}_1024.txt"))
        assert a.shape == (1024, 2)
        arrays[f"{op}_1024"] = a
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({len(arrays)} arrays)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
