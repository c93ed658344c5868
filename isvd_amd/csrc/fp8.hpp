// FP8 sketch: A stored as OCP e4m3 (e4m3fn: bias 7, largest finite value 448, no infinity)
// with one FP32 scale per 128 x 128 block, multiplied on the gfx950 block-scaled FP8 matrix
// cores (v_mfma_scale_f32_16x16x128_f8f6f4), FP32 accumulation.  No counterpart in the
// reference; the FP64 sketch it replaces is sketchBlockRow
// (src/lib/libisvd/core/stage/@x@_sketch.h:136-139).
//
// Storage (device pointers, P = 1):
//   a8      m x n codes, row-major: row i at a8 + i * lda8, lda8 >= n, lda8 % 16 == 0, the
//           pointer 16-byte aligned;
//   scale   ceil(m/128) x ceil(n/128) FP32, row-major (block (bi, bj) at scale[bi * lds + bj]);
//           A[i, j] = decode(a8[i, j]) * scale[i / 128, j / 128].  Every scale must be a power
//           of two in [2^-127, 2^127]: the matrix cores apply it exactly as an E8M0 exponent.
//   omega8  Omega (n x Nl), column c at omega8 + c * ldo (the K index contiguous), ldo >= n,
//           ldo % 16 == 0, unscaled e4m3 codes.
#pragma once

#include "stages.hpp"

namespace isvd {

constexpr i64 kF8Block = 128;  // edge of a scale block

// a8 / scale from an FP32 matrix (ordera 'C': a[i + j * lda], 'R': a[i * lda + j]).  Each
// block's scale is the smallest power of two s with max|A_block| <= 448 s (1 for an all-zero
// block); codes are A / s rounded to nearest even.
void quantize_e4m3(char ordera, i64 m, i64 n, const float *a, i64 lda, uint8_t *a8, i64 lda8,
                   float *scale, i64 lds);

// omega8 := e4m3 codes of the FP32 Gaussian stream of gaussian_omega<float> (Omega[j, c] =
// that stream's Omega^T[c, j]), rounded to nearest even
void gaussian_omega_e(i64 n, i64 Nl, uint8_t *omega8, i64 ldo, unsigned long long seed);

// Y^T (Nl x m, FP32, ldyst) := Omega^T A^T from the e4m3 operands.  Fails (before any product)
// if a scale is not a power of two in [2^-127, 2^127].
void stage_sketch_e(const Dims &d, const uint8_t *a8, i64 lda8, const float *scale, i64 lds,
                    const uint8_t *omega8, i64 ldo, float *yst, i64 ldyst);

}  // namespace isvd
