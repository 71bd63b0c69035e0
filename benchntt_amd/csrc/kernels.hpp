// kernels.hpp -- launch helpers shared by the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "u128.hpp"

namespace bntt {

enum BlasOp : int { OP_VVADD = 0, OP_VVSUB = 1, OP_VVMUL = 2, OP_AXPY = 3 };

// Elementwise ops over n residues (device pointers).  x is the axpy scalar.
hipError_t launch_blas(int op, U128 *c, const U128 *a, const U128 *b, U128 x, size_t n, const Mod &md,
                       hipStream_t stream);

// Forward constant-geometry NTT of `count` transforms of size 2^log2n, stored
// contiguously, all with twiddle table tw (2^log2n entries).  tmp must hold
// count * 2^log2n residues when log2n > kLdsMaxLog2 (multi-pass sizes); it may
// be null otherwise, and with it a lone transform of 2^kSpreadMinLog2 ..
// 2^kLdsMaxLog2 points takes two passes.  x and y may alias only when
// log2n <= kLdsMaxLog2.  With
// `done` (one transform of at most 2^kLdsMaxLog2 points: a single workgroup)
// the kernel stores `seq` to *done, at system scope, after its outputs.
hipError_t launch_ntt_uniform(U128 *y, const U128 *x, const U128 *tw, int log2n, size_t count, const Mod &md,
                              U128 *tmp, hipStream_t stream, uint32_t *done = nullptr, uint32_t seq = 0);

constexpr int kLdsMaxLog2 = 12;   // transforms up to 4096 points run in one LDS-resident pass
constexpr int kSpreadMinLog2 = 11;  // ... except a lone one from 2^11 points, given a tmp: two passes
constexpr int kMinLog2 = 1;
constexpr int kMaxLog2 = 24;

}  // namespace bntt
