"""benchntt_amd -- MI355X (gfx950) implementation of benchNTT's hot path.

128-bit modular forward NTT (constant-geometry network of ntt_scalar.c:582-627)
and vvadd / vvsub / vvmul / axpy mod q (blas_scalar.c:578-608) as hand-written
HIP kernels behind the C ABI of include/bntt.h (libbntt.so, built in-tree by
``python benchntt_amd/build.py``).  Importing this package loads the library
and raises ImportError when it is missing; there is no CPU fallback.

Modules: ops (the reference's names on device tensors or host arrays), batch
(uneven batches: work-queue plans, graph replay, tune table), polymul
(negacyclic products with continuous batching), fixtures (text fixtures, word
order, root tables, seeded residues), autotune, workloads, residues, config.
"""
from . import config  # noqa: F401
from ._lib import BnttError, U128, load as _load_lib  # noqa: F401

_load_lib()

from .ops import axpy, ntt, tmp_elements, vector_add, vector_mul, vector_sub  # noqa: E402,F401

__all__ = ["ntt", "vector_add", "vector_sub", "vector_mul", "axpy", "tmp_elements", "BnttError", "config"]
