#!/bin/bash
# build_ref.sh -- compile the reference's own C kernels, where they lie under
# /root/reference/src, into oracle/_ref/*.so (git-ignored) for use as a second
# checker and as the CPU baseline.  TEST INFRASTRUCTURE ONLY.
#
# Nothing from the reference is copied into this repository: each object is
# compiled straight from the reference's source file plus oracle/ref_glue.c
# (pointer-only wrappers of our own) and the reference's config.c (its fixture
# loaders, which main() references).  main() is renamed so that several of
# these objects can be loaded into one process.
#
# NTT sizes: the reference fixes N at compile time with `#define stages K`
# (ntt_scalar.c:21 = 18, ntt_avx2.c:18 = 10, ntt_avx512.c:20 = 10) and sweeps N
# by rewriting that one line with sed before each compile
# (benchmark/run_ntt.sh:14).  The same substitution is applied here on a pipe
# into the compiler, so no edited copy of any file is written anywhere.
#
# Flags follow the reference Makefile (-O3 -w, Makefile:8-9) with a portable
# -march in place of -march=native, because the objects are built here and run
# on another host: x86-64-v3 (AVX2) and x86-64-v4 (AVX-512F/BW/CD/DQ/VL).
set -euo pipefail

REF=${BNTT_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/src/ntt_scalar.c" ]; then
    echo "build_ref: no reference sources at $REF/src; skipping"
    exit 0
fi
mkdir -p "$OUT"
CC=${CC:-gcc}
COMMON=(-O3 -w -fPIC -fno-semantic-interposition -shared -I"$REF/src" -Dmain=bntt_ref_unused_main)
GLUE="$HERE/ref_glue.c"
CFG="$REF/src/config.c"

build_blas() {  # variant macro march
    local src=$1 def=$2 march=$3 out=$4
    $CC "${COMMON[@]}" -march=$march -D$def "$REF/src/$src" "$CFG" "$GLUE" -o "$OUT/$out"
}

build_ntt() {   # file macro march stages out
    local src=$1 def=$2 march=$3 stages=$4 out=$5
    local obj="$OUT/.${out%.so}.o"
    sed "s/^#define stages [0-9]*/#define stages ${stages}/" "$REF/src/$src" |
        $CC -O3 -w -fPIC -fno-semantic-interposition -c -march=$march -I"$REF/src" -Dmain=bntt_ref_unused_main -x c - -o "$obj"
    $CC "${COMMON[@]}" -march=$march -D$def -DBNTT_REF_STAGES=$stages "$obj" "$CFG" "$GLUE" -o "$OUT/$out"
    rm -f "$obj"
}

build_blas blas_scalar.c REF_BLAS_SCALAR x86-64-v3 ref_blas_scalar.so &
build_blas blas_avx2.c REF_BLAS_AVX2 x86-64-v3 ref_blas_avx2.so &
build_blas blas_avx512.c REF_BLAS_AVX512 x86-64-v4 ref_blas_avx512.so &
wait
# scalar NTT at every size the reference sweep covers, plus the small fixture sizes
for L in 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 20; do
    build_ntt ntt_scalar.c REF_NTT_SCALAR x86-64-v3 $L ref_ntt_scalar_L$L.so &
done
wait
for L in 4 6 8 10 12 14 16; do
    build_ntt ntt_avx2.c REF_NTT_AVX2 x86-64-v3 $L ref_ntt_avx2_L$L.so &
    build_ntt ntt_avx512.c REF_NTT_AVX512 x86-64-v4 $L ref_ntt_avx512_L$L.so &
done
wait
n=$(ls "$OUT"/*.so | wc -l)
if [ "$n" -ne 34 ]; then
    echo "build_ref: expected 34 objects, found $n" >&2
    exit 1
fi
echo "build_ref: built $n objects into $OUT"
