// kernels.hpp -- launch helpers shared by the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "u128.hpp"

namespace bntt {

enum BlasOp : int { OP_VVADD = 0, OP_VVSUB = 1, OP_VVMUL = 2, OP_AXPY = 3 };

// Workgroups go to the XCDs round-robin by index (workgroup b to XCD b mod 8).
// Chunk order of a grid of one-chunk workgroups: block b takes chunk
// (b mod 8) * (nb8 / 8) + b / 8 of the first nb8 = nblocks - nblocks mod 8
// chunks, so each XCD works through one contiguous eighth of the data; the
// last nblocks mod 8 blocks keep their own chunks.  Any such order is a
// permutation of the chunks, so only speed depends on where the workgroups
// actually run.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ size_t xcd_chunk(size_t blk, size_t nblocks)
{
    const size_t nb8 = nblocks - nblocks % kXcds;
    return blk < nb8 ? (blk % kXcds) * (nb8 / kXcds) + blk / kXcds : blk;
}

// Elementwise ops over n residues (device pointers).  x is the axpy scalar.
// With `done` (host-array calls) each of the launch's blas_blocks(n)
// workgroups w posts seq to done[w] once its outputs are visible system-wide.
hipError_t launch_blas(int op, U128 *c, const U128 *a, const U128 *b, U128 x, size_t n, const Mod &md,
                       hipStream_t stream, uint32_t *done = nullptr, uint32_t seq = 0);
size_t blas_blocks(size_t n);

// Forward constant-geometry NTT of `count` transforms of size 2^log2n, stored
// contiguously, all with twiddle table tw (2^log2n entries).  tmp must hold
// count * 2^log2n residues when log2n > kLdsMaxLog2 (multi-pass sizes); it may
// be null otherwise.  Given for one transform of 2^11 .. 2^13 points, it must
// hold 2^log2n + 16 residues (the intermediate and the grid-barrier words),
// and the transform splits over several workgroups (or, in a launch being
// captured into a graph, takes one workgroup or two launches).  x and y may
// alias only when log2n <= kLdsMaxLog2.  With `done` (one transform of at most
// 2^kLoneSplitMaxLog2 points, one launch) each of the launch's *done_words
// workgroups w stores `seq` to done[w], at system scope, after its outputs; a
// launch that would need more than done_cap words is refused
// (hipErrorInvalidValue) before anything runs.
hipError_t launch_ntt_uniform(U128 *y, const U128 *x, const U128 *tw, int log2n, size_t count, const Mod &md,
                              U128 *tmp, hipStream_t stream, uint32_t *done = nullptr, uint32_t seq = 0,
                              int *done_words = nullptr, int done_cap = 0);

constexpr int kLdsMaxLog2 = 12;   // transforms up to 4096 points run in one LDS-resident pass
constexpr int kLoneSplitMinLog2 = 11;  // a lone transform from 2^11 points given a tmp splits over workgroups
constexpr int kLoneSplitMaxLog2 = 13;  // ... up to 2^13 points (beyond, the barrier over 32 workgroups costs more than it saves)
constexpr int kSpreadMinLog2 = 11;  // ... except a lone one from 2^11 points given a tmp under capture: two launches
// Workgroups of a lone transform of 2^L points given a tmp (1: one workgroup).
__host__ __device__ constexpr int lone_parts(int L)
{
    return L == 11 ? 4 : L == 12 ? 8 : L == 13 ? 16 : 1;
}
constexpr int kMinLog2 = 1;
constexpr int kMaxLog2 = 24;

}  // namespace bntt
