// BF16 products with FP32 accumulation for the mixed-precision pipeline (the hand-written
// matrix-core kernel of bf16_gemm.hip), plus the precision conversions around them.  FP32 data enter a product as materialized BF16 splits
// x = h0 + h1 (+ h2), each term the round-to-nearest-even BF16 of the remainder: two terms for
// the "split bf16 x3" Gram (hi*hi + hi*lo + lo*hi), three terms (exact to FP32) against exact
// BF16 data for the projection.
#pragma once

#include "common.hpp"

namespace isvd {

// C (m x n, FP32, ldc) := alpha op(A) op(B) + beta C for BF16 A, B (column-major, op 'N' / 'T')
// on the BF16 matrix cores with FP32 accumulation
void gemm_bf16(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A, i64 lda,
               const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc);

// dst (rows x cols, ldd) := src (rows x cols, lds), converted
template <typename Tout, typename Tin>
void convert(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd);

// BF16 bits (round to nearest even) of an FP32 array
void f32_to_bf16(i64 count, const float *src, uint16_t *dst);

// dst (parts*rows x cols, ldd), rows p*rows .. p*rows + rows - 1 := BF16 term p of src
// (rows x cols, lds): x = h0 + h1 (+ h2), each term the nearest-even BF16 of the remainder
void split_bf16(i64 rows, i64 cols, int parts, const float *src, i64 lds, uint16_t *dst,
                i64 ldd);

// dst (rows x cols, ldd) := the sum of the `parts` row blocks of src (parts*rows x cols)
void sum_blocks(i64 rows, i64 cols, int parts, const float *src, i64 lds, float *dst, i64 ldd);

// M (n x n) := sym(HH) + LH + LH^T from C = [H^T H ; L^T H] (2n x n)
void gram_combine(i64 n, const float *C, i64 ldc, float *M, i64 ldm);

}  // namespace isvd
