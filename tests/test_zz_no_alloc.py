"""GPU check of bntt_ntt_fwd_dev_noalloc, the forward NTT that allocates
nothing (its multi-pass intermediates live in a fixed device array of the
library): called through the C ABI, its outputs must equal bntt_ntt_fwd_dev's
(with a caller-allocated tmp) bit for bit, on consecutive calls that reuse the
array with a smaller input the second time, and sizes beyond the array are
rejected."""
import ctypes

import numpy as np
import pytest

from benchntt_amd import _lib
from benchntt_amd import config as C
from benchntt_amd._lib import U128
from benchntt_amd.residues import uniform
from conftest import from_dev, to_dev

pytestmark = pytest.mark.gpu

QN, MUN = C.Q_NTT, C.MU_NTT
MAXR = _lib.BNTT_NOALLOC_MAX_RESIDUES


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def with_tmp(torch, dx, dtw, L, count):
    dy = torch.empty_like(dx)
    need = _lib.load().bntt_ntt_tmp_bytes(L, count) // 16
    tmp = torch.empty((max(need, 1), 2), dtype=torch.int64, device=dx.device)
    _lib.call("bntt_ntt_fwd_dev", dy.data_ptr(), dx.data_ptr(), dtw.data_ptr(), L, count,
              ctypes.byref(U128.of(QN)), ctypes.byref(U128.of(MUN)), tmp.data_ptr() if need else None,
              _stream(torch))
    torch.cuda.synchronize()
    return from_dev(dy)


def noalloc(torch, dx, dtw, L, count):
    dy = torch.empty_like(dx)
    torch.cuda.synchronize()
    _lib.call("bntt_ntt_fwd_dev_noalloc", dy.data_ptr(), dx.data_ptr(), dtw.data_ptr(), L, count,
              ctypes.byref(U128.of(QN)), ctypes.byref(U128.of(MUN)), _stream(torch))
    torch.cuda.synchronize()
    return from_dev(dy)


def test_noalloc_matches_tmp_entry(gpu):
    """Two calls in a row: the whole array (one 2^20-point transform, two
    passes), then a smaller input (5 transforms of 2^14); then a one-pass
    size, which uses no scratch; then one transform of 2^13 and of 2^11
    points, split over several workgroups with their barrier words in the
    array (twice at 2^13: the second launch meets the first one's words)."""
    for L, count, seed in ((20, 1, 1), (14, 5, 2), (10, 7, 3), (13, 1, 4), (13, 1, 5), (11, 1, 6)):
        n = 1 << L
        print(f"noalloc L={L} count={count}", flush=True)
        x, tw = uniform(count * n, QN, 100 + seed), uniform(n, QN, 200 + seed)
        dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
        want = with_tmp(gpu, dx, dtw, L, count)
        got = noalloc(gpu, dx, dtw, L, count)
        assert np.array_equal(got, want), (L, count)


def test_noalloc_three_pass_size(gpu):
    """Sizes above 2^16 points take three passes (two through the scratch
    array is not the pattern then: y, scratch, y)."""
    L, count = 17, 4
    n = 1 << L
    x, tw = uniform(count * n, QN, 31), uniform(n, QN, 32)
    dx, dtw = to_dev(gpu, x), to_dev(gpu, tw)
    assert np.array_equal(noalloc(gpu, dx, dtw, L, count), with_tmp(gpu, dx, dtw, L, count))


def test_noalloc_rejects_beyond_scratch(gpu):
    """count * 2^log2n above the array is an argument error, checked before
    anything is launched (so the pointers are never dereferenced)."""
    dx = gpu.empty((16, 2), dtype=gpu.int64, device="cuda")
    for L, count in ((20, 2), (21, 1), (13, MAXR // (1 << 13) + 1)):
        with pytest.raises(_lib.BnttError) as ei:
            _lib.call("bntt_ntt_fwd_dev_noalloc", dx.data_ptr(), dx.data_ptr(), dx.data_ptr(), L, count,
                      ctypes.byref(U128.of(QN)), ctypes.byref(U128.of(MUN)), _stream(gpu))
        assert ei.value.code == 1 and "scratch" in str(ei.value)
