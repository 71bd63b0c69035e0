"""Seeded synthetic workloads of the benchmark configurations (host side).

uneven_pairs(): the uneven batch -- `count` forward NTTs in pairs, each pair of
one size N = 2^L with L uniform on {lo..hi} (log-uniform N), both operands
uniform residues in [0, q); one shared seeded twiddle table per size; the
pointwise product of each pair's outputs follows its two transforms.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .residues import uniform


@dataclass
class UnevenBatch:
    log2n: np.ndarray        # per transform
    offsets: np.ndarray      # element offset of each transform in the flat buffers
    total: int               # elements over all transforms
    x: np.ndarray            # (total, 2) uint64 inputs
    tables: dict             # log2n -> (2^log2n, 2) uint64 twiddle table
    pair_offsets: np.ndarray  # element offset of each pair's product in the product buffer
    pair_total: int


def uneven_pairs(count: int, q: int, seed: int, lo: int = 4, hi: int = 16, with_data: bool = True) -> UnevenBatch:
    if count % 2:
        raise ValueError("transforms come in pairs: count must be even")
    rng = np.random.Generator(np.random.PCG64(seed))
    pair_l = rng.integers(lo, hi + 1, size=count // 2)
    log2n = np.repeat(pair_l, 2)
    sizes = np.left_shift(1, log2n.astype(np.int64))
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    psizes = np.left_shift(1, pair_l.astype(np.int64))
    pair_offsets = np.concatenate([[0], np.cumsum(psizes)[:-1]])
    x = uniform(total, q, seed + 1) if with_data else None
    tables = {int(l): uniform(1 << int(l), q, seed + 1000 + int(l)) for l in np.unique(log2n)}
    return UnevenBatch(log2n, offsets, total, x, tables, pair_offsets, int(psizes.sum()))
