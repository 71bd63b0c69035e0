import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def fixtures():
    """The reference's golden vectors (tests/golden/ref_fixtures.npz)."""
    with np.load(os.path.join(GOLDEN, "ref_fixtures.npz")) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.load()


@pytest.fixture(scope="session")
def gpu():
    """torch with a usable GPU and the HIP library loaded; fails (does not
    skip) when a gpu-marked test runs without them."""
    import torch

    assert torch.cuda.is_available(), "gpu tests need a GPU"
    import benchntt_amd  # noqa: F401  (raises if libbntt.so is missing)

    return torch


def to_dev(torch, a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


def from_dev(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64).reshape(-1, 2)
