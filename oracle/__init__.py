"""CPU oracle for the benchNTT hot path.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the product:
``benchntt_amd`` does not import it and fails loudly when its HIP library is
missing instead of falling back to anything here.

``oracle.c`` restates the reference's scalar algorithm (ntt_scalar.c:49-199,
582-627; blas_scalar.c:578-608); ``ref.py`` drives the reference's own kernels
compiled by ``build_ref.sh`` into ``_ref/`` when that build exists.
"""
from .oracle import (  # noqa: F401
    build,
    load,
    negacyclic_tables,
    Oracle,
    root_table,
    two_adic_root,
)
