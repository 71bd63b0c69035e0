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
# Two flavours of every object:
#   portable  gcc -O3 -w with -march=x86-64-v3 (AVX2) / x86-64-v4 (AVX-512F/BW/
#             CD/DQ/VL) in place of the Makefile's default -march=native
#             (Makefile:8-9), because the objects are built here and run on
#             another host; these run anywhere the ISA exists (tests here)
#   amd       the reference's own AMD configuration, AMD_CC = clang with
#             AMD_CFLAGS = -march=znver4 -O3 -w (Makefile:4-5), with the ROCm
#             clang; these are the CPU baseline on hosts with znver4's ISA.
#             clang 22 makes an incompatible pointer type an error by default;
#             the AVX2 files pass uint64_t * to _mm256_storeu_si256, which the
#             clang of the reference's time only warned about (silenced by -w),
#             so that error is turned back into a warning for them
# Sizes: L = 4..20 for every ISA (benchmark/run_ntt.sh:13 sweeps 8..20).
set -euo pipefail

REF=${BNTT_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/src/ntt_scalar.c" ]; then
    echo "build_ref: no reference sources at $REF/src; skipping"
    exit 0
fi
mkdir -p "$OUT"
GCC=${CC:-gcc}
CLANG=${BNTT_REF_CLANG:-/opt/rocm/llvm/bin/clang}
GLUE="$HERE/ref_glue.c"
CFG="$REF/src/config.c"
COMMON=(-O3 -w -fPIC -fno-semantic-interposition -I"$REF/src" -Dmain=bntt_ref_unused_main)

# flags of one compiler: clang's pointer-type error back to a warning (above)
cflags() { if [ "$1" = "$CLANG" ]; then echo "-Wno-error=incompatible-pointer-types"; fi; }

build_blas() {  # cc march file macro out
    local cc=$1 march=$2 src=$3 def=$4 out=$5
    $cc "${COMMON[@]}" $(cflags $cc) -shared -march=$march -D$def "$REF/src/$src" "$CFG" "$GLUE" -o "$OUT/$out"
}

build_ntt() {   # cc march file macro stages out
    local cc=$1 march=$2 src=$3 def=$4 stages=$5 out=$6
    local obj="$OUT/.${out%.so}.o"
    sed "s/^#define stages [0-9]*/#define stages ${stages}/" "$REF/src/$src" |
        $cc "${COMMON[@]}" $(cflags $cc) -c -march=$march -x c - -o "$obj"
    $cc "${COMMON[@]}" $(cflags $cc) -shared -march=$march -D$def -DBNTT_REF_STAGES=$stages "$obj" "$CFG" "$GLUE" \
        -o "$OUT/$out"
    rm -f "$obj"
}

flavours="portable"
if [ -x "$CLANG" ] && "$CLANG" -march=znver4 -E -x c /dev/null -o /dev/null 2>/dev/null; then
    flavours="portable amd"
fi
jobs=0
throttle() { jobs=$((jobs + 1)); if [ $jobs -ge ${BNTT_REF_JOBS:-8} ]; then wait; jobs=0; fi; }
for fl in $flavours; do
    if [ $fl = portable ]; then
        cc=$GCC; m2=x86-64-v3; m5=x86-64-v4; sfx=""
    else
        cc=$CLANG; m2=znver4; m5=znver4; sfx="_amd"
    fi
    build_blas $cc $m2 blas_scalar.c REF_BLAS_SCALAR ref_blas_scalar$sfx.so & throttle
    build_blas $cc $m2 blas_avx2.c REF_BLAS_AVX2 ref_blas_avx2$sfx.so & throttle
    build_blas $cc $m5 blas_avx512.c REF_BLAS_AVX512 ref_blas_avx512$sfx.so & throttle
    for L in $(seq 4 20); do
        build_ntt $cc $m2 ntt_scalar.c REF_NTT_SCALAR $L ref_ntt_scalar${sfx}_L$L.so & throttle
        build_ntt $cc $m2 ntt_avx2.c REF_NTT_AVX2 $L ref_ntt_avx2${sfx}_L$L.so & throttle
        build_ntt $cc $m5 ntt_avx512.c REF_NTT_AVX512 $L ref_ntt_avx512${sfx}_L$L.so & throttle
    done
done
wait
want=$((54 * $(echo $flavours | wc -w)))
n=$(ls "$OUT"/*.so | wc -l)
if [ "$n" -lt "$want" ]; then
    echo "build_ref: expected $want objects, found $n" >&2
    exit 1
fi
echo "build_ref: built $n objects ($flavours) into $OUT"
