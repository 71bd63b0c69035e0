"""Seeded synthetic residues and array conversions.

A residue array is a numpy ``uint64`` array of shape ``(n, 2)`` (column 0 the
low word, column 1 the high word) -- the memory layout of unsigned __int128 --
or a torch tensor with the same bytes (``int64``, shape ``(n, 2)``).
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def uniform(n: int, q: int, seed: int) -> np.ndarray:
    """n residues uniform in [0, q) from numpy's PCG64 stream `seed`:
    uniform draws below 2^bitlen(q), with the rare draws >= q redrawn."""
    bits = q.bit_length()
    if not 65 <= bits <= 128:
        raise ValueError("q must have 65 to 128 bits")
    hbits = bits - 64
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((n, 2), dtype=np.uint64)
    out[:, 0] = rng.integers(0, 1 << 64, size=n, dtype=np.uint64, endpoint=False)
    out[:, 1] = rng.integers(0, 1 << hbits, size=n, dtype=np.uint64, endpoint=False)
    qh, ql = np.uint64(q >> 64), np.uint64(q & MASK64)
    while True:
        bad = (out[:, 1] > qh) | ((out[:, 1] == qh) & (out[:, 0] >= ql))
        k = int(bad.sum())
        if k == 0:
            return out
        out[bad, 0] = rng.integers(0, 1 << 64, size=k, dtype=np.uint64, endpoint=False)
        out[bad, 1] = rng.integers(0, 1 << hbits, size=k, dtype=np.uint64, endpoint=False)


def to_ints(a) -> list[int]:
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in a]


def from_ints(vals) -> np.ndarray:
    vals = list(vals)
    out = np.empty((len(vals), 2), dtype=np.uint64)
    out[:, 0] = [v & MASK64 for v in vals]
    out[:, 1] = [(v >> 64) & MASK64 for v in vals]
    return out
